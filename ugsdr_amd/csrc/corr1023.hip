// Correlation kernel for the C/A-code lengths N = 1023 * M (M = 4, 8, 12, 16:
// 1 ms at 4.092, 8.184, 12.276 and 16.368 MHz, or 4 ms at 4.092 MHz).
//
// The inverse FFT of one row block runs as three Stockham stages,
// N = 31 * 33 * M, with one radix-31 butterfly per thread (N / 31 threads,
// rounded up to whole waves):
//   stage 1 (radix 31, no twiddles): the thread loads its 31 inputs
//           x[i] = X[(i + shift) mod N] * conj(C)[i] straight from global
//           memory into registers, so the product never round-trips LDS;
//   stage 2 (radix 33 = 3 x 11 prime factor): twiddles w_1023^(r k) come
//           from a 31 x 33 table, one LDS read each;
//   stage 3 (radix M): butterfly j owns outputs m = j + q * 1023, q < M.
//           The thread keeps the same butterflies for every block, so |.|
//           (or the complex fold) accumulates in registers and the fold to a
//           surface of L = 1023 * QL samples is q -> q mod QL, in-thread.
// One transform is two LDS round trips and four workgroup barriers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "acq.hpp"
#include "corr_common.hpp"
#include "ctx.hpp"
#include "dev_fft.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

constexpr int kC1023 = 1023;

template <int M>
struct C1023Shape {
    static constexpr int N = kC1023 * M;
    static constexpr int B1 = 33 * M;   // stage-1 butterflies
    static constexpr int B2 = 31 * M;   // stage-2 butterflies
    static constexpr int T = (B1 + 63) / 64 * 64;  // whole waves
    static constexpr int NSLOT = (kC1023 + T - 1) / T;  // stage-3 butterflies per thread
    // LDS: data, stage-2 table [31][33], stage-3 table [1023], reduction scratch
    static constexpr size_t lds_bytes = (size_t)(N + 2 * kC1023) * sizeof(float2) + 128 * sizeof(double);
};

// w^r for r = 0..R-1 from w^1 by a shallow product tree (depth <= 4 products
// for R <= 16, so the powers stay within a few ulp of the table value)
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 w1, float2 (&w)[R]) {
    w[0] = make_float2(1.f, 0.f);
    if constexpr (R > 1) w[1] = w1;
    if constexpr (R > 2) w[2] = cmul(w1, w1);
    if constexpr (R > 3) w[3] = cmul(w[2], w1);
    if constexpr (R > 4) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w1);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = cmul(w[4], w[4]);
#pragma unroll
        for (int r = 9; r < R; ++r) w[r] = cmul(w[8], w[r - 8]);
    }
}

// LDS of the three-stage kernels: data, stage-2 table, stage-3 table, reduction scratch
template <int M>
struct C1023Lds {
    float2 *buf, *tw2, *tw3;
    double* red;
    __device__ C1023Lds(unsigned char* smem, int tid) {
        constexpr int N = C1023Shape<M>::N, T = C1023Shape<M>::T;
        buf = reinterpret_cast<float2*>(smem);
        tw2 = buf + N;         // [k * 33 + r] = exp(+2 pi i r k / 1023)
        tw3 = tw2 + kC1023;    // [j] = exp(+2 pi i j / N)
        red = reinterpret_cast<double*>(tw3 + kC1023);
        for (int i = tid; i < kC1023; i += T) {
            const int k = i / 33, r = i - k * 33;
            double s, c;
            sincospi(2.0 * (double)((r * k) % kC1023) / (double)kC1023, &s, &c);
            tw2[i] = make_float2((float)c, (float)s);
            sincospi(2.0 * (double)i / (double)N, &s, &c);
            tw3[i] = make_float2((float)c, (float)s);
        }
    }
};

// Blocks k0 .. k1-1 of one row into the thread's surface samples
// S[s][u] = surface[tid + s T + u 1023].
template <int M, int QL>
__device__ __forceinline__ void corr1023_blocks(float (&S)[C1023Shape<M>::NSLOT][QL], const C1023Lds<M>& Ld,
                                                const ReqDesc& q, int cls, int sh, gptr<float2> C, int k0, int k1,
                                                int fold_mode, int tid) {
    using S_ = C1023Shape<M>;
    constexpr int N = S_::N, T = S_::T, B1 = S_::B1, B2 = S_::B2, NSLOT = S_::NSLOT;
    constexpr int FOLD = M / QL;
    float2* buf = Ld.buf;
    const float2* tw2 = Ld.tw2;
    const float2* tw3 = Ld.tw3;
    for (int k = k0; k < k1; ++k) {
        const int t = opaque(tid);
        const gptr<float2> Xk = as_global(q.X + ((int64_t)k * q.F + cls) * N);
        // ---- stage 1: radix 31 from global memory
        // (every thread computes, idle ones on a clamped index, so the
        // registers are defined on all paths; only the stores are guarded)
        {
            const bool act = t < B1;
            const int j = act ? t : 0;
            float2 v[31];
            // gathers in two halves: the second half's addresses wait for the
            // first half's products, bounding the registers loads hold
            int jb = j;
#pragma unroll
            for (int r = 0; r < 31; ++r) {
                if (r == 16) jb = after(jb, v[15].x);
                const int i = jb + r * B1;
                int src = i + sh;
                src -= src >= N ? N : 0;
                v[r] = cmul(Xk[src], C[i]);
            }
            DftOdd<31, true>::prep(v);
            __syncthreads();
            if (act) DftOdd<31, true>::store(v, buf + j * 31, 1);
        }
        __syncthreads();
        // ---- stage 2: radix 33 (3 x 11), Ns = 31
        {
            const bool act = t < B2;
            const int j = act ? t : 0;
            const int kk = j % 31;
            float2 u[33];
#pragma unroll
            for (int r = 0; r < 33; ++r) u[r] = buf[j + r * B2];
            // twiddles in groups of 8 reads, each group after the last one's products
            int tb = kk * 33;
#pragma unroll
            for (int r = 1; r < 33; ++r) {
                if (r > 1 && (r - 1) % 8 == 0) tb = after(tb, u[r - 1].x);
                u[r] = cmul(u[r], tw2[tb + r]);
            }
            Pfa33<true>::prep(u);
            __syncthreads();
            if (act) {
                float2* dst = buf + (j - kk) * 33 + kk;
                Pfa33<true>::store(u, [&](int qq, float2 x) { dst[qq * 31] = x; });
            }
        }
        __syncthreads();
        // ---- stage 3: radix M, Ns = 1023, accumulate |.| per owned output
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            // one slot at a time: the next slot's loads wait for this one's sums
            if (s > 0) order_after(S[s - 1][0]);
            const int j = opaque(t) + s * T;
            if (j < kC1023) {
                float2 w[M], tw[M];
#pragma unroll
                for (int r = 0; r < M; ++r) w[r] = buf[j + r * kC1023];
                twiddle_powers<M>(tw3[j], tw);
#pragma unroll
                for (int r = 1; r < M; ++r) w[r] = cmul(w[r], tw[r]);
                Dft<M, true>::run(w);
                if (fold_mode == 0) {
#pragma unroll
                    for (int u = 0; u < QL; ++u) {
                        float acc = 0.f;
#pragma unroll
                        for (int f = 0; f < FOLD; ++f) {
                            const float2 z = w[u + f * QL];
                            acc += cabs_hw(z);
                        }
                        S[s][u] += acc;
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < QL; ++u) {
                        float2 z = w[u];
#pragma unroll
                        for (int f = 1; f < FOLD; ++f) z = cadd(z, w[u + f * QL]);
                        S[s][u] += cabs_hw(z);
                    }
                }
            }
        }
    }

}

// surface samples of a finished row -> row statistics (and the surface)
template <int M, int QL>
__device__ __forceinline__ void corr1023_stats(const float (&S)[C1023Shape<M>::NSLOT][QL], double* red,
                                               const ReqDesc& q, int local, float inv_n, int tid) {
    using S_ = C1023Shape<M>;
    constexpr int T = S_::T, NSLOT = S_::NSLOT, L = kC1023 * QL;
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int j = tid + s * T;
        if (j < kC1023) {
#pragma unroll
            for (int u = 0; u < QL; ++u) {
                top2_merge(v1, i1, v2, S[s][u], j + u * kC1023, -1.0f);
                sum += (double)S[s][u];
            }
        }
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double qv = 0.0;
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int j = tid + s * T;
        if (j < kC1023) {
#pragma unroll
            for (int u = 0; u < QL; ++u) {
                const double d = (double)S[s][u] - mean;
                qv += d * d;
            }
        }
    }
    qv = block_sum<T>(qv, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, qv, L, tid, q.rowstats + local, inv_n);
    if (q.surfaces) {
        float* o = q.surfaces + (int64_t)local * L;
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            const int j = tid + s * T;
            if (j < kC1023) {
#pragma unroll
                for (int u = 0; u < QL; ++u) o[j + u * kC1023] = S[s][u] * inv_n;
            }
        }
    }
}

// Tail row `tail` from the stream-K shares' partial surfaces, summed in
// share order (a fixed order, so results do not depend on timing), then its
// statistics.  G is the stream-K kernel's grid.  S has room for L floats.
template <int T>
__device__ __forceinline__ void corr1023_merge_row(const CorrArgs& a, int tail, int G, float* S, double* red,
                                                   int tid) {
    const int L = a.L;
    const int row = a.full_rows + tail;
    const ReqDesc q = a.reqs[find_req(a, row)];
    const int local = row - q.row_start;
    const int ur0 = q.tail_unit_start + (local - q.tail_row0) * q.K;
    const int wa = sk_owner(a.units, G, ur0), wb = sk_owner(a.units, G, ur0 + q.K - 1);
    // parts in share order; L is a multiple of 4 (1023 * QL, QL in 4, 8, 12, 16)
    const float* part = reinterpret_cast<const float*>(a.scratch);
    float4* S4 = reinterpret_cast<float4*>(S);
    const int L4 = L / 4;
    for (int w = wa; w <= wb; ++w) {
        const int64_t base = wa == wb ? ((int64_t)2 * G + tail) * L
                                      : ((int64_t)w * 2 + (sk_begin(a.units, G, w) >= ur0 ? 0 : 1)) * L;
        const float4* src = reinterpret_cast<const float4*>(part + base);
        for (int j = tid; j < L4; j += T) {
            const float4 v = src[j];
            if (w == wa) {
                S4[j] = v;
            } else {
                float4 o = S4[j];
                o.x += v.x;
                o.y += v.y;
                o.z += v.z;
                o.w += v.w;
                S4[j] = o;
            }
        }
    }
    __syncthreads();
    row_stats_glob<T>(S, L, red, tid, q.rowstats + local, a.inv_n);
    if (q.surfaces) {
        float* o = q.surfaces + (int64_t)local * L;
        for (int j = tid; j < L; j += T) o[j] = S[j] * a.inv_n;
    }
}

// One row per workgroup: every block, then the statistics.  The first
// a.n_merge workgroups instead finish the tail rows from the stream-K
// partial surfaces (launched before this kernel), so that short work runs
// beside the first round of rows rather than after the last.
template <int M, int QL>
__global__ __launch_bounds__(C1023Shape<M>::T, 3) void corr1023(CorrArgs a) {
    using S_ = C1023Shape<M>;
    constexpr int NSLOT = S_::NSLOT;
    static_assert(M % QL == 0, "surface length must divide the transform");
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x < a.n_merge) {
        float* S = reinterpret_cast<float*>(smem);
        double* red = reinterpret_cast<double*>(S + (a.L + 1) / 2 * 2);
        corr1023_merge_row<S_::T>(a, blockIdx.x, a.sk_grid, S, red, tid);
        return;
    }
    const C1023Lds<M> Ld(smem, tid);
    const int row = blockIdx.x - a.n_merge;
    const ReqDesc q = a.reqs[find_req(a, row)];
    const int local = row - q.row_start;
    const int p = local / q.D, di = local - p * q.D;
    float S[NSLOT][QL];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s)
#pragma unroll
        for (int u = 0; u < QL; ++u) S[s][u] = 0.f;
    corr1023_blocks<M, QL>(S, Ld, q, q.cls[di], q.shift[di], as_global(q.specs[p]), 0, q.K, a.fold_mode, tid);
    corr1023_stats<M, QL>(S, Ld.red, q, local, a.inv_n, tid);
}

// Stream-K over the tail rows (fewer than one round of workgroups): the
// tail's (row, block) units are split evenly over all workgroups.  Workgroup
// w's share is [begin(w), begin(w + 1)) of the tail units.  A row the share
// holds whole goes to the row's own slot (scratch after the 2G share slots);
// a row cut by the share's start or end goes to share slot 0 or 1 (only the
// first and the last row of a share can be cut).  corr1023_merge finishes
// the rows.
template <int M, int QL>
__global__ __launch_bounds__(C1023Shape<M>::T, 3) void corr1023_sk(CorrArgs a) {
    using S_ = C1023Shape<M>;
    constexpr int T = S_::T, NSLOT = S_::NSLOT, L = kC1023 * QL;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const C1023Lds<M> Ld(smem, tid);
    const int w = blockIdx.x;
    const int u_begin = sk_begin(a.units, gridDim.x, w), u_end = sk_begin(a.units, gridDim.x, w + 1);
    for (int u = u_begin; u < u_end;) {
        const ReqDesc q = a.reqs[find_req_unit(a, u)];
        const int lu = u - q.tail_unit_start;
        const int local = q.tail_row0 + lu / q.K, k0 = lu % q.K;
        const int k1 = min(q.K, k0 + (u_end - u));
        const int p = local / q.D, di = local - p * q.D;
        float S[NSLOT][QL];
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
#pragma unroll
            for (int v = 0; v < QL; ++v) S[s][v] = 0.f;
        corr1023_blocks<M, QL>(S, Ld, q, q.cls[di], q.shift[di], as_global(q.specs[p]), k0, k1, a.fold_mode, tid);
        float* part = reinterpret_cast<float*>(a.scratch);
        const bool whole = k0 == 0 && k1 == q.K;
        float* dst = whole ? part + ((int64_t)2 * gridDim.x + (q.row_start + local - a.full_rows)) * L
                           : part + ((int64_t)w * 2 + (u == u_begin ? 0 : 1)) * L;
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            const int j = tid + s * T;
            if (j < kC1023) {
#pragma unroll
                for (int v = 0; v < QL; ++v) dst[j + v * kC1023] = S[s][v];
            }
        }
        u += k1 - k0;
    }
}

template <class K>
static hipError_t lds_attr(K kernel, size_t bytes) {
    return set_lds_limit(reinterpret_cast<const void*>(kernel), bytes);
}

template <int M, int QL>
static int launch_1023(const CorrArgs& c, hipStream_t st) {
    const size_t lds = C1023Shape<M>::lds_bytes;
    constexpr int T = C1023Shape<M>::T;
    if (c.sk_grid > 0) {
        UGSDR_HIP_TRY(lds_attr(corr1023_sk<M, QL>, lds));
        hipLaunchKernelGGL((corr1023_sk<M, QL>), dim3(c.sk_grid), dim3(T), lds, st, c);
    }
    if (c.full_rows + c.n_merge > 0) {
        UGSDR_HIP_TRY(lds_attr(corr1023<M, QL>, lds));
        hipLaunchKernelGGL((corr1023<M, QL>), dim3(c.full_rows + c.n_merge), dim3(T), lds, st, c);
    }
    return UGSDR_OK;
}

int corr1023_threads(int64_t N) {
    if (N % kC1023) return 0;
    const int64_t M = N / kC1023;
    switch (M) {
        case 4: return C1023Shape<4>::T;
        case 8: return C1023Shape<8>::T;
        case 12: return C1023Shape<12>::T;
        case 16: return C1023Shape<16>::T;
        default: return 0;
    }
}

bool corr1023_supports(int64_t N, int64_t L) {
    if (!corr1023_threads(N) || L % kC1023 || N % L) return false;
    const int64_t M = N / kC1023, QL = L / kC1023;
    return QL == M || QL == 4;
}

// workgroups of the three-stage kernel one CU holds (LDS bound)
static int per_cu(int M) {
    size_t lds = 0;
    switch (M) {
        case 4: lds = C1023Shape<4>::lds_bytes; break;
        case 8: lds = C1023Shape<8>::lds_bytes; break;
        case 12: lds = C1023Shape<12>::lds_bytes; break;
        default: lds = C1023Shape<16>::lds_bytes; break;
    }
    return std::max(1, (int)(160 * 1024 / lds));
}

int corr1023_grid(int64_t N, int cus) { return cus * per_cu((int)(N / kC1023)); }

int launch_corr1023(const CorrArgs& c, int cus, DevBuf& scratch, hipStream_t st) {
    const int M = c.N / kC1023, QL = c.L / kC1023;
    if (!corr1023_supports(c.N, c.L)) {
        set_error("corr1023: unsupported length pair");
        return UGSDR_ERR_UNSUPPORTED;
    }
    const int tail_rows = c.rows - c.full_rows;
    const int G_sk = tail_rows > 0 ? std::min(corr1023_grid(c.N, cus), c.units) : 0;
    CorrArgs a = c;
    a.sk_grid = G_sk;
    a.n_merge = G_sk > 0 ? tail_rows : 0;
    if (G_sk > 0) {
        UGSDR_HIP_TRY(scratch.reserve(((size_t)G_sk * 2 + tail_rows) * c.L * 4));
        a.scratch = scratch.as<float2>();
    }
    switch (M * 100 + QL) {
        case 404: return launch_1023<4, 4>(a, st);
        case 808: return launch_1023<8, 8>(a, st);
        case 804: return launch_1023<8, 4>(a, st);
        case 1212: return launch_1023<12, 12>(a, st);
        case 1204: return launch_1023<12, 4>(a, st);
        case 1616: return launch_1023<16, 16>(a, st);
        case 1604: return launch_1023<16, 4>(a, st);
        default: set_error("corr1023: unsupported length pair"); return UGSDR_ERR_UNSUPPORTED;
    }
}

}  // namespace ugsdr_hip
