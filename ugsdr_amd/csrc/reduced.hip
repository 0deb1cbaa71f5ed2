// Reduced-precision acquisition (FP16 / BF16 storage, matrix-core DFTs).
//
//   replicas : conj(DFT) stored as 16-bit pairs scaled by 1/sqrt(N);
//   classes  : the FP32 class spectra X[k][c] of each search are packed to
//              16-bit pairs with a power-of-two scale per block (k, c), so
//              every block uses the full 16-bit range;
//   corr     : N = N1 * 341 four-step inverse DFT per row and block, every
//              stage (N1 = Ra * Rb, 341 = 11 * 31) a matrix-core product with
//              FP32 accumulation; the DFT matrices carry 1/R, so values stay
//              O(1) and the result is already 1/N scaled; |.| accumulates in
//              FP32 with the block scales undone, into a workgroup-private
//              surface in global memory, then the FP32 row statistics.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "acq.hpp"
#include "corr_common.hpp"
#include "ctx.hpp"
#include "dev_mfma.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

__global__ void to_half2_kernel(const float2* in, __half2* out, int64_t n, float scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __floats2half2_rn(in[i].x * scale, in[i].y * scale);
}
__global__ void to_bf162_kernel(const float2* in, __hip_bfloat162* out, int64_t n, float scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __hip_bfloat162(__float2bfloat16(in[i].x * scale), __float2bfloat16(in[i].y * scale));
}
__global__ void from_half2_kernel(const __half2* in, float2* out, int64_t n, float inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = __half22float2(in[i]);
        out[i] = make_float2(v.x * inv, v.y * inv);
    }
}
__global__ void from_bf162_kernel(const __hip_bfloat162* in, float2* out, int64_t n, float inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = __bfloat1622float2(in[i]);
        out[i] = make_float2(v.x * inv, v.y * inv);
    }
}

int replica_reduce(ugsdr_hip_ctx* ctx, Replica& rep, int precision) {
    // Parseval: sum |C_k|^2 = N * sum c^2 = N^2 for +-1 replicas, so the RMS
    // bin magnitude is sqrt(N); scaling by 1/sqrt(N) centres it at 1.
    const float scale = (float)(1.0 / std::sqrt((double)rep.n_fft));
    void* dst = nullptr;
    UGSDR_HIP_TRY(hipMalloc(&dst, (size_t)rep.n_fft * 4));
    const int blocks = (int)std::min<int64_t>((rep.n_fft + 255) / 256, 4096);
    if (precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(to_half2_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.spec), static_cast<__half2*>(dst), rep.n_fft, scale);
    else
        hipLaunchKernelGGL(to_bf162_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.spec), static_cast<__hip_bfloat162*>(dst), rep.n_fft,
                           scale);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    UGSDR_HIP_TRY(hipFree(rep.spec));
    rep.spec = dst;
    rep.precision = precision;
    rep.scale = scale;
    return UGSDR_OK;
}

int replica_read(ugsdr_hip_ctx* ctx, const Replica& rep, float* out) {
    if (rep.precision == UGSDR_PREC_FP32) {
        UGSDR_HIP_TRY(hipMemcpyAsync(out, rep.spec, (size_t)rep.n_fft * 8, hipMemcpyDeviceToHost, ctx->stream));
        UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return UGSDR_OK;
    }
    UGSDR_HIP_TRY(ctx->aux.reserve((size_t)rep.n_fft * 8));
    const int blocks = (int)std::min<int64_t>((rep.n_fft + 255) / 256, 4096);
    if (rep.precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(from_half2_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const __half2*>(rep.spec), ctx->aux.as<float2>(), rep.n_fft, 1.0f / rep.scale);
    else
        hipLaunchKernelGGL(from_bf162_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const __hip_bfloat162*>(rep.spec), ctx->aux.as<float2>(), rep.n_fft,
                           1.0f / rep.scale);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, ctx->aux.ptr, (size_t)rep.n_fft * 8, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

// ------------------------------------------------------------------ class blocks
__global__ void zero_u32_kernel(unsigned* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// max |X| of block blockIdx.y (magnitudes are non-negative, so their bit
// patterns order like unsigned integers): one atomic per workgroup, so the
// workgroups of a block do not queue on its word
__global__ __launch_bounds__(256) void block_absmax_kernel(const float2* X, int64_t N, unsigned* bmax) {
    __shared__ float wm[4];
    const float2* x = X + (int64_t)blockIdx.y * N;
    float m = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, sqrtf(fmaf(x[i].x, x[i].x, x[i].y * x[i].y)));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(bmax + blockIdx.y, __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

// 2^-e with max |X| = f 2^e, f in [0.5, 1): the block lands in [0.5, 1)
__device__ __forceinline__ int block_exponent(unsigned bits) {
    const float m = __uint_as_float(bits);
    int e = 0;
    if (m > 0.f) frexpf(m, &e);
    return e;
}

template <class H>
__global__ void block_pack_kernel(const float2* X, int64_t N, const unsigned* bmax, uint32_t* Xr, float* inv) {
    const int b = blockIdx.y;
    const int e = block_exponent(bmax[b]);
    const float s = ldexpf(1.f, -e);
    const float2* x = X + (int64_t)b * N;
    uint32_t* o = Xr + (int64_t)b * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = pack16<H>(make_float2(x[i].x * s, x[i].y * s));
    if (blockIdx.x == 0 && threadIdx.x == 0) inv[b] = ldexpf(1.f, e);
}

// ------------------------------------------------------------------ correlation
constexpr int kRedCap = 16384;   // 32-bit pairs per LDS buffer (two buffers)
constexpr int kRedN2 = 341;      // second four-step factor, 11 * 31
constexpr int kRedGP = 344;      // words per phase-1 output row in global memory
                                 // (>= 341 + 3, a multiple of 4: 16-byte stores)

// scratch words of one workgroup: max_k phase-1 slices of N / 341 rows,
// then the FP32 surface
__host__ __device__ constexpr int64_t red_wg_words(int64_t N, int max_k) {
    return ((int64_t)max_k * (N / kRedN2) * kRedGP + 3) / 4 * 4;
}

// smallest v >= x with v % mod == rem
constexpr int pad_to(int x, int mod, int rem) { return x + ((rem - x) % mod + mod) % mod; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }

// LDS layouts of the four stages (SeqLayout rows: stride 1 mod 32 plus the
// bank skew).  A stage's input element (m, jj) of a sequence (m < R, jj <
// PER butterflies) sits at base(seq) + m * MS + jj with MS odd; the stage
// writing that input puts its output (jjo, q) there as (m = jjo, jj = q).
//   phase 1 (columns, length N1 = Ra Rb): in (m < Ra, jj < Rb) [MS = Rb,
//     LD1] holds element n1 at base + n1; mid (m < Rb, jj < Ra) [SX1 =
//     Ra | 1, LDB1]; the second stage's output k1 = jj + q Ra goes straight
//     to global memory, row k1 of the block's slice (GP words a row);
//   phase 2 (rows, 341 = 31 * 11): in (m < 31, jj < 11) [11, LD2] holds
//     element n2 at base + n2; mid (m < 11, jj < 31) [31, LD2]; k2 = jj +
//     31 q; the surface rows stage through both buffers as (k2, seq) at
//     k2 LS + seq + skew(seq), LS % 32 == 31 (31 LS == 1 mod 32: a tile's
//     writes are conflict free).
// Batches hold whole tiles of 16 sequences.
template <int R1a, int R1b>
struct RedShape {
    static constexpr int N1 = R1a * R1b, N2 = kRedN2;
    static constexpr int SK = SeqLayout<1>::SKEW;
    static constexpr int SX1 = R1a | 1;
    static_assert(R1b % 2 == 1, "phase-1 first radix needs an odd row");
    static constexpr int LD1 = pad_to(N1 + SK, 32, 1), LDB1 = pad_to(R1b * SX1 + SK, 32, 1);
    static constexpr int ROW1 = cmax(N1, R1b * SX1), LDM1 = cmax(LD1, LDB1);
    static constexpr int CB = cmin(((kRedCap - SK - ROW1) / LDM1 + 1) / 16 * 16, (N2 + 15) / 16 * 16);
    static constexpr int GP = kRedGP;
    static_assert(GP >= N2 + 3 && GP % 4 == 0, "phase-1 rows take 16-byte stores of 4 columns");
    static constexpr int LD2 = pad_to(N2 + SK, 32, 1);
    static constexpr int RB = ((kRedCap - SK - N2) / LD2 + 1) / 16 * 16;
    static constexpr int LS = pad_to(RB + SK, 32, 31);
    static_assert(CB >= 16 && RB >= 16, "an LDS buffer holds less than one tile");
    static_assert(N2 * LS <= 2 * kRedCap, "surface staging exceeds the LDS buffers");
    // passes of the last stage per wave (its outputs stay in registers)
    template <int T, int GPP>
    static constexpr int passes2() {
        return ((RB + 15) / 16 * 31 + GPP * (T / 64) - 1) / (GPP * (T / 64));
    }
    static size_t lds_bytes(int N) {
        return (size_t)2 * kRedCap * 4 +
               (size_t)(MfmaShape<R1a>::ENTRIES + MfmaShape<R1b>::ENTRIES + MfmaShape<31>::ENTRIES +
                        MfmaShape<11>::ENTRIES) * 16 +
               (size_t)(2 * N1 + kRedN2 + kTwLo + (N + kTwLo - 1) / kTwLo) * 8 + 128 * 8;
    }
};

// One workgroup per row (grid-strided).  Phase 1 runs the column DFTs of
// every block of the row (two matrix-core stages; the inter-stage twiddle
// and the four-step twiddle are applied to FP32 stage outputs) into a
// workgroup slice of global memory; phase 2 runs the row DFTs per batch of
// rows over every block, the last stage accumulating |.| of its FP32
// outputs in registers.  Each finished surface value goes straight into the
// thread's share of the row statistics (top two with the first index of the
// maximum, FP64 sum and sum of squares), so the surface itself is written
// only when the caller asks for it.
template <class H, int R1a, int R1b, int T, int GPP>
__global__ __launch_bounds__(T, T / 256) void corr_mfma(CorrArgs a) {
    using L = RedShape<R1a, R1b>;
    using LA = SeqLayout<L::LD1>;
    using LB = SeqLayout<L::LDB1>;
    using L2 = SeqLayout<L::LD2>;
    constexpr int NW = T / 64, N1 = L::N1, N2 = L::N2, GP = L::GP;
    constexpr int P2 = L::template passes2<T, GPP>();
    constexpr int RPP1 = T / L::CB;   // product pass: column tc, rows tr + i RPP1
    constexpr int RPP2 = T / N2;      // row copy: column tc, rows tr + i RPP2
    static_assert(RPP1 >= 1 && RPP2 >= 1, "workgroup narrower than a batch row");
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* bufA = reinterpret_cast<uint32_t*>(smem);
    uint32_t* bufB = bufA + kRedCap;
    uint4* bt1a = reinterpret_cast<uint4*>(bufB + kRedCap);
    uint4* bt1b = bt1a + MfmaShape<R1a>::ENTRIES;
    uint4* bt2a = bt1b + MfmaShape<R1b>::ENTRIES;
    uint4* bt2b = bt2a + MfmaShape<31>::ENTRIES;
    float2* tw1 = reinterpret_cast<float2*>(bt2b + MfmaShape<11>::ENTRIES);
    float2* twk = tw1 + N1;
    float2* tw2 = twk + N1;
    float2* twlo = tw2 + N2;
    float2* twhi = twlo + kTwLo;
    const int N = a.N;
    double* red = reinterpret_cast<double*>(twhi + (N + kTwLo - 1) / kTwLo);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    mfma_btab_build<R1a, H>(bt1a, true, 1.f / R1a, tid, T);
    mfma_btab_build<R1b, H>(bt1b, true, 1.f / R1b, tid, T);
    mfma_btab_build<31, H>(bt2a, true, 1.f / 31, tid, T);
    mfma_btab_build<11, H>(bt2b, true, 1.f / 11, tid, T);
    tw_table_build(tw1, N1, true, tid, T);
    tw_table_build(tw2, N2, true, tid, T);
    tw_build(twlo, twhi, N, tid, T);
    for (int j = tid; j < N1; j += T) {   // exp(+2 pi i j / N), the four-step step
        double sn, cs;
        sincospi(2.0 * (double)j / (double)N, &sn, &cs);
        twk[j] = make_float2((float)cs, (float)sn);
    }
    __syncthreads();
    // workgroup scratch: the phase-1 output of every block of a row (maxK
    // slices of N1 rows k1 of GP words, columns n2 < N2 used)
    uint32_t* Gall = reinterpret_cast<uint32_t*>(a.scratch) + (int64_t)blockIdx.x * red_wg_words(N, a.max_k);
    for (int row = blockIdx.x; row < a.rows; row += gridDim.x) {
        const ReqDesc q = a.reqs[find_req(a, row)];
        const int local = row - q.row_start;
        // the thread's share of the row statistics
        float v1 = -1.0f, v2 = -1.0f;
        int i1 = 0x7fffffff;
        double sum = 0.0, sumsq = 0.0;
        float* surf = q.surfaces ? q.surfaces + (int64_t)local * N : nullptr;
        const int p = local / q.D, di = local - p * q.D;
        const int cls = q.cls[di], sh = q.shift[di];
        const gptr<uint32_t> C = as_global(reinterpret_cast<const uint32_t*>(q.specs[p]));
        for (int k = 0; k < q.K; ++k) {
            const int t0 = opaque(tid);
            const gptr<uint32_t> X =
                as_global(reinterpret_cast<const uint32_t*>(q.X) + ((int64_t)k * q.F + cls) * N);
            const gwptr<uint32_t> G = as_global_w(Gall + (int64_t)k * N1 * GP);
            const int tc = t0 % L::CB, tr = t0 / L::CB;
            for (int c0 = 0; c0 < N2; c0 += L::CB) {
                const int nb = min(L::CB, N2 - c0);
                // x (shifted class block) times the replica spectrum: column
                // tc of the batch, element n1 at base(tc) + n1.  The previous
                // batch's second stage reads only the other buffer.
                if (tr < RPP1 && tc < nb) {
                    uint32_t* dst = bufA + LA::base(tc);
                    const int nc = c0 + tc;
                    for (int n1 = tr; n1 < N1; n1 += 4 * RPP1) {
                        uint32_t xv[4], cv[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int m = n1 + u * RPP1;
                            const int n = nc + N2 * (m < N1 ? m : 0);
                            int src = n + sh;
                            src -= src >= N ? N : 0;
                            xv[u] = X[src];
                            cv[u] = C[n];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (n1 + u * RPP1 < N1) dst[n1 + u * RPP1] = cmul16<H>(xv[u], cv[u]);
                    }
                }
                __syncthreads();
                // radix Ra; output (jj, q) times w_N1^(jj q) is the next input
                mfma_tiles<R1a, H, NW, GPP, 0, R1b, R1b, LA>(
                    bufA, nb, bt1a, lane, wave, [&](int, int, int sb, int jj, const f32x4* Dre, const f32x4* Dim) {
                        const int s0 = sb * 16 + 4 * lg;
                        if (li < R1a && s0 < nb) {
                            const float2 w = tw1[jj * li];
                            uint32_t* o = bufB + LB::base(s0) + jj * L::SX1 + li;
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (s0 + r < nb)
                                    o[r * L::LDB1] = pack16<H>(cmul(make_float2(Dre[0][r], Dim[0][r]), w));
                        }
                    });
                __syncthreads();
                // radix Rb; output k1 = jj + q Ra of columns n2 = c0 + s0 + r
                // times the four-step twiddle w_N^(n2 k1), four columns in one
                // 16-byte store (past the last column it lands in row padding)
                mfma_tiles<R1b, H, NW, GPP, 0, R1a, L::SX1, LB>(
                    bufB, nb, bt1b, lane, wave, [&](int, int, int sb, int jj, const f32x4* Dre, const f32x4* Dim) {
                        const int s0 = sb * 16 + 4 * lg;
                        if (li < R1b && s0 < nb) {
                            const int k1 = jj + li * R1a;
                            float2 w = tw_get<true>(twlo, twhi, (c0 + s0) * k1);
                            const float2 st = twk[k1];
                            uint32_t v[4];
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                v[r] = pack16<H>(cmul(make_float2(Dre[0][r], Dim[0][r]), w));
                                w = cmul(w, st);
                            }
                            *(gwptr<uint4>)(G + k1 * GP + c0 + s0) = make_uint4(v[0], v[1], v[2], v[3]);
                        }
                    });
            }
        }
        __syncthreads();   // the slices are complete and visible to every wave
        // phase 2 per batch of rows k1 = r0 + t over every block
        const int tc2 = opaque(tid) % N2, tr2 = opaque(tid) / N2;
        for (int r0 = 0; r0 < N1; r0 += L::RB) {
            const int nb = min(L::RB, N1 - r0);
            float acc[P2][GPP][4];
#pragma unroll
            for (int i0 = 0; i0 < P2; ++i0)
#pragma unroll
                for (int i1 = 0; i1 < GPP; ++i1)
#pragma unroll
                    for (int i2 = 0; i2 < 4; ++i2) acc[i0][i1][i2] = 0.f;
            for (int k = 0; k < q.K; ++k) {
                // rows r0 .. r0 + nb - 1 of slice k into base(t) + n2 (the
                // previous block's last stage reads only the other buffer)
                const gptr<uint32_t> g = as_global(Gall + ((int64_t)k * N1 + r0) * GP + tc2);
                if (tr2 < RPP2)
                    for (int t = tr2; t < nb; t += 4 * RPP2) {
                        uint32_t v[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int tt = t + u * RPP2;
                            v[u] = g[(tt < nb ? tt : t) * GP];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int tt = t + u * RPP2;
                            if (tt < nb) bufA[L2::base(tt) + tc2] = v[u];
                        }
                    }
                __syncthreads();
                mfma_tiles<31, H, NW, GPP, 0, 11, 11, L2>(
                    bufA, nb, bt2a, lane, wave, [&](int, int, int sb, int jj, const f32x4* Dre, const f32x4* Dim) {
                        const int s0 = sb * 16 + 4 * lg;
                        if (s0 < nb) {
#pragma unroll
                            for (int qt = 0; qt < 2; ++qt) {
                                const int qq = qt * 16 + li;
                                if (qq < 31) {
                                    const float2 w = tw2[jj * qq];
                                    uint32_t* o = bufB + L2::base(s0) + jj * 31 + qq;
#pragma unroll
                                    for (int r = 0; r < 4; ++r)
                                        if (s0 + r < nb)
                                            o[r * L::LD2] = pack16<H>(cmul(make_float2(Dre[qt][r], Dim[qt][r]), w));
                                }
                            }
                        }
                    });
                __syncthreads();
                const float sc = q.block_inv[k * q.F + cls] * q.spec_inv;
                mfma_tiles<11, H, NW, GPP, P2, 31, 31, L2>(
                    bufB, nb, bt2b, lane, wave, [&](int pass, int gi, int, int, const f32x4* Dre, const f32x4* Dim) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc[pass][gi][r] += cabs_hw(make_float2(Dre[0][r], Dim[0][r])) * sc;
                    });
            }
            // the batch's surface rows through LDS as (k2, seq), then to S in
            // natural order m = r0 + seq + N1 k2
            __syncthreads();
            float* stage = reinterpret_cast<float*>(bufA);
#pragma unroll
            for (int pass = 0; pass < P2; ++pass)
#pragma unroll
                for (int gi = 0; gi < GPP; ++gi) {
                    const int t = mfma_slot_tile<NW, GPP, 31>(pass, gi, nb, wave);
                    const int sb = t / 31, jj = t - sb * 31;
                    const int s0 = sb * 16 + 4 * lg;
                    if (t >= 0 && li < 11) {
                        float* o = stage + (jj + 31 * li) * L::LS + s0 + L2::skew(s0);
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (s0 + r < nb) o[r] = acc[pass][gi][r];
                    }
                }
            __syncthreads();
            walk_cols<4>(
                opaque(tid), T, nb, N2, [&](int t, int k2) { return stage[k2 * L::LS + t + L2::skew(t)]; },
                [&](int t, int k2, float v) {
                    const int m = r0 + t + N1 * k2;
                    top2_merge(v1, i1, v2, v, m, -1.0f);
                    sum += (double)v;
                    sumsq += (double)v * (double)v;
                    if (surf) surf[m] = v;
                });
            __syncthreads();
        }
        // sum of squared deviations from the FP64 sums: the surface is
        // positive with a spread comparable to its mean, so the difference
        // keeps all but a digit or two of FP64's precision
        block_top2<T>(v1, i1, v2, red, tid);
        sum = block_sum<T>(sum, red + 64, tid);
        sumsq = block_sum<T>(sumsq, red + 64, tid);
        const double mean = sum / (double)N;
        const double dev = sumsq - sum * mean;
        stats_write<T>(v1, i1, v2, mean, dev > 0.0 ? dev : 0.0, N, tid, q.rowstats + local, 1.0f);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ upsampled code
// Reduced-precision upsampled-code search: every replica is one period of
// Lc = N1 * 341 chips held for U samples each (the BOC(1,1) 4 ms code: 8184
// half-chips at 10 samples), so a row (PRN x Doppler bin) is, per phase
// p < U, the Lc-point inverse transform of Z_p conj(C) over every block,
// with Z formed once per (bin, block) by the Z-prep kernel as 16-bit pairs
// with a power-of-two scale per (bin, block).  N1, 31 and 11 are pairwise
// coprime, so each phase transform runs as a prime-factor (Good-Thomas)
// product Lc = N1 x 31 x 11 with no twiddle between stages, held in LDS by
// one workgroup per row:
//   columns: element k' = (341 n1 + N1 n2) mod Lc sits in column n2, row
//            n1; Z and the code spectrum are stored in that order (storage
//            word n1 x 341 + c, c = b + 11 a for n2 = (11 a + 31 b) mod 341),
//            so a tile's 16 columns are 16 consecutive words.  The 16-bit
//            products Z_p conj(C) go from global memory straight into the
//            matrix-core A operands, one radix-N1 stage, packed into LDS as
//            N1 rows of 341 (row k1, word c);
//   rows:    radix 31 over a (11 butterflies b), radix 11 over b (31
//            butterflies k_a), both on the matrix cores; |.| of the last
//            stage's FP32 outputs with the scales undone is summed over the
//            blocks in registers.
// Output (k1, k_a, k_b) is transform sample m = (k1 E1 + k2 E2) mod Lc with
// k2 = (187 k_a + 155 k_b) mod 341 (the CRT idempotents), row sample
// tau = U m + p; a phase's sums go into the thread's share of the row
// statistics before the next.
constexpr int kUpmLD = pad_to(kRedN2 + SeqLayout<1>::SKEW, 32, 1);

// modular inverse of a mod m (gcd 1), for the prime-factor index maps
__host__ __device__ constexpr int mod_inverse(int a, int m) {
    for (int x = 1; x < m; ++x)
        if ((a % m) * x % m == 1) return x;
    return 0;
}

// storage word of element k' (< N1 x 341) and the natural index of storage word s
template <int N1>
struct UpmPfa {
    static constexpr int N2 = kRedN2, Lc = N1 * N2;
    static constexpr int I1 = mod_inverse(N2 % N1, N1), I2 = mod_inverse(N1, N2);
    // CRT idempotents of Lc = N1 x 341: E1 = 1 mod N1, 0 mod 341; E2 the reverse
    static constexpr int E1 = N2 * I1 % Lc, E2 = N1 * I2 % Lc;
    static_assert(I1 > 0 && I2 > 0, "N1 and 341 must be coprime");
    __host__ __device__ static int pos(int k) {
        const int n1 = (k % N1) * I1 % N1, n2 = (k % N2) * I2 % N2;
        const int a = (n2 % 31) * 17 % 31, b = (n2 % 11) * 5 % 11;
        return n1 * N2 + b + 11 * a;
    }
    __host__ __device__ static int natural(int s) {
        const int n1 = s / N2, c = s - n1 * N2, a = c / 11, b = c - 11 * a;
        const int n2 = (11 * a + 31 * b) % N2;
        return (N2 * n1 + N1 * n2) % Lc;
    }
};
static_assert(UpmPfa<24>::E1 == 1705 && UpmPfa<24>::E2 == 6480, "CRT idempotents of 8184");
static_assert(17 * 11 % 31 == 1 && 5 * 31 % 11 == 1 && 187 % 31 == 1 && 187 % 11 == 0 && 155 % 31 == 0 &&
              155 % 11 == 1, "prime-factor maps of 341 = 31 x 11");

// in place: each block of Lc 16-bit words from natural order into storage
// order, through LDS (coalesced reads and writes)
template <int N1>
__global__ __launch_bounds__(1024) void upm_permute_kernel(uint32_t* z) {
    using P = UpmPfa<N1>;
    __shared__ uint32_t buf[P::Lc];
    uint32_t* b = z + (int64_t)blockIdx.x * P::Lc;
    for (int i = threadIdx.x; i < P::Lc; i += blockDim.x) buf[i] = b[i];
    __syncthreads();
    for (int i = threadIdx.x; i < P::Lc; i += blockDim.x) b[i] = buf[P::natural(i)];
}

int upm_permute(uint32_t* blocks, int64_t nblocks, int64_t code_len, hipStream_t st) {
    if (nblocks <= 0) return UGSDR_OK;
    switch (code_len / kRedN2) {
        case 12: hipLaunchKernelGGL(upm_permute_kernel<12>, dim3((unsigned)nblocks), dim3(1024), 0, st, blocks); break;
        case 24: hipLaunchKernelGGL(upm_permute_kernel<24>, dim3((unsigned)nblocks), dim3(1024), 0, st, blocks); break;
        default: set_error("prime-factor storage order: unsupported code length"); return UGSDR_ERR_UNSUPPORTED;
    }
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

template <int N1>
struct UpmShape {
    static constexpr int N2 = kRedN2, Lc = N1 * N2;
    using Lay = SeqLayout<kUpmLD>;
    // words of one buffer of N1 rows (base(N1 - 1) + N2, rounded up to 4)
    static constexpr int BUF = ((N1 - 1) * kUpmLD + Lay::SKEW + N2 + 3) / 4 * 4;
    static constexpr int CT = (N2 + 15) / 16;   // column tiles of 16 columns
    static constexpr size_t lds_bytes =
        (size_t)2 * BUF * 4 +
        (size_t)(MfmaShape<N1>::ENTRIES + MfmaShape<31>::ENTRIES + MfmaShape<11>::ENTRIES) * 16 + 128 * 8;
};

template <class H, int N1, int T>
__global__ __launch_bounds__(T, T / 256) void corr_upm(CorrArgs a) {
    using S = UpmShape<N1>;
    using Lay = typename S::Lay;
    constexpr int NW = T / 64, N2 = S::N2, Lc = S::Lc;
    constexpr int GPP = T >= 1024 ? 1 : 2;
    constexpr int SB = (N1 + 15) / 16;   // sequence tiles of the row stages
    constexpr int P2 = (SB * 31 + GPP * NW - 1) / (GPP * NW);
    constexpr int KS1 = MfmaShape<N1>::KS, QT1 = MfmaShape<N1>::QT;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* bufR = reinterpret_cast<uint32_t*>(smem);
    uint32_t* bufM = bufR + S::BUF;
    uint4* bt1 = reinterpret_cast<uint4*>(bufM + S::BUF);
    uint4* bt2a = bt1 + MfmaShape<N1>::ENTRIES;
    uint4* bt2b = bt2a + MfmaShape<31>::ENTRIES;
    double* red = reinterpret_cast<double*>(bt2b + MfmaShape<11>::ENTRIES);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    mfma_btab_build<N1, H>(bt1, true, 1.f / N1, tid, T);
    mfma_btab_build<31, H>(bt2a, true, 1.f / 31, tid, T);
    mfma_btab_build<11, H>(bt2b, true, 1.f / 11, tid, T);
    __syncthreads();
    const int U = a.ucode;
    const int row = up_row_of(blockIdx.x, a.rows, a.row_group);
    const ReqDesc q = a.reqs[find_req(a, row)];
    const RowRef rr = row_ref(q, row - q.row_start, true);
    const gptr<uint32_t> C = as_global(reinterpret_cast<const uint32_t*>(q.specs[rr.p]));
    float* surf = q.surfaces ? q.surfaces + (int64_t)rr.slot * a.L : nullptr;
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0, sumsq = 0.0;
    for (int p = 0; p < U; ++p) {
        float acc[P2][GPP][4];
#pragma unroll
        for (int i0 = 0; i0 < P2; ++i0)
#pragma unroll
            for (int i2 = 0; i2 < GPP; ++i2)
#pragma unroll
                for (int i3 = 0; i3 < 4; ++i3) acc[i0][i2][i3] = 0.f;
        for (int k = 0; k < q.K; ++k) {
            const int dk = rr.di * q.K + k;
            const gptr<uint32_t> Z = as_global(reinterpret_cast<const uint32_t*>(q.X) + ((int64_t)dk * U + p) * Lc);
            // ---- columns: radix N1 from global memory into LDS rows.  The
            // previous block's first row stage, the last reader of bufR,
            // ended before the barrier that preceded its second stage.
            {
                for (int t = wave; t < S::CT; t += NW) {
                    const int c = t * 16 + li;
                    const int cc = c < N2 ? c : 0;
                    typename H::v8 A[KS1];
#pragma unroll
                    for (int ks = 0; ks < KS1; ++ks) {
                        uint32_t zv[4], cv[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int m = ks * 16 + 4 * lg + e;
                            const int n = cc + N2 * (m < N1 ? m : 0);
                            zv[e] = Z[n];
                            cv[e] = C[n];
                        }
                        const int m0 = ks * 16 + 4 * lg;
                        uint4 av;
                        av.x = m0 + 0 < N1 ? cmul16<H>(zv[0], cv[0]) : 0u;
                        av.y = m0 + 1 < N1 ? cmul16<H>(zv[1], cv[1]) : 0u;
                        av.z = m0 + 2 < N1 ? cmul16<H>(zv[2], cv[2]) : 0u;
                        av.w = m0 + 3 < N1 ? cmul16<H>(zv[3], cv[3]) : 0u;
                        A[ks] = __builtin_bit_cast(typename H::v8, av);
                    }
                    // B operands from LDS per output tile (fewer live registers)
                    f32x4 Dre[QT1], Dim[QT1];
#pragma unroll
                    for (int qt = 0; qt < QT1; ++qt) {
                        Dre[qt] = f32x4{0.f, 0.f, 0.f, 0.f};
                        Dim[qt] = Dre[qt];
#pragma unroll
                        for (int ks = 0; ks < KS1; ++ks) {
                            const auto bre = __builtin_bit_cast(typename H::v8, bt1[((ks * QT1 + qt) * 2 + 0) * 64 + lane]);
                            const auto bim = __builtin_bit_cast(typename H::v8, bt1[((ks * QT1 + qt) * 2 + 1) * 64 + lane]);
                            Dre[qt] = H::mfma(A[ks], bre, Dre[qt]);
                            Dim[qt] = H::mfma(A[ks], bim, Dim[qt]);
                        }
                    }
                    // lane (li, lg): output k1 = qt 16 + li of storage columns s0 .. s0 + 3
                    const int s0 = t * 16 + 4 * lg;
#pragma unroll
                    for (int qt = 0; qt < QT1; ++qt) {
                        const int k1 = qt * 16 + li;
                        if (k1 < N1 && s0 < N2) {
                            uint32_t* o = bufR + Lay::base(k1) + s0;
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (s0 + r < N2) o[r] = pack16<H>(make_float2(Dre[qt][r], Dim[qt][r]));
                        }
                    }
                }
            }
            __syncthreads();
            // ---- rows, radix 31 over a (11 butterflies b a row): output
            // (b, k_a) at word 31 b + k_a, no twiddle
            mfma_tiles<31, H, NW, GPP, 0, 11, 11, Lay>(
                bufR, N1, bt2a, lane, wave, [&](int, int, int sb, int jj, const f32x4* Dre, const f32x4* Dim) {
                    const int s0 = sb * 16 + 4 * lg;
                    if (s0 < N1) {
#pragma unroll
                        for (int qt = 0; qt < 2; ++qt) {
                            const int qq = qt * 16 + li;
                            if (qq < 31) {
                                uint32_t* o = bufM + Lay::base(s0) + jj * 31 + qq;
#pragma unroll
                                for (int r = 0; r < 4; ++r)
                                    if (s0 + r < N1) o[r * kUpmLD] = pack16<H>(make_float2(Dre[qt][r], Dim[qt][r]));
                            }
                        }
                    }
                });
            __syncthreads();
            // ---- rows, radix 11 over b (31 butterflies k_a a row): output
            // k_b, |.| with the Z, replica and 1/U scales undone.  The next
            // block's first row stage, which overwrites bufM, comes after the
            // barrier that ends its column stage.
            const float sc = q.block_inv[dk] * q.spec_inv;
            mfma_tiles<11, H, NW, GPP, P2, 31, 31, Lay>(
                bufM, N1, bt2b, lane, wave, [&](int pass, int gi, int, int, const f32x4* Dre, const f32x4* Dim) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[pass][gi][r] += cabs_hw(make_float2(Dre[0][r], Dim[0][r])) * sc;
                });
        }
        // phase p's samples into the thread's share of the row statistics
#pragma unroll
        for (int pass = 0; pass < P2; ++pass)
#pragma unroll
            for (int gi = 0; gi < GPP; ++gi) {
                const int t = mfma_slot_tile<NW, GPP, 31>(pass, gi, N1, wave);
                if (t >= 0 && li < 11) {
                    const int sb = t / 31, jj = t - sb * 31;
                    const int k2 = (187 * jj + 155 * li) % N2;   // CRT of (k_a, k_b) = (jj, li)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int k1 = sb * 16 + 4 * lg + r;
                        if (k1 < N1) {
                            const int m = (k1 * UpmPfa<N1>::E1 + k2 * UpmPfa<N1>::E2) % Lc;
                            const int tau = U * m + p;
                            const float v = acc[pass][gi][r];
                            top2_merge(v1, i1, v2, v, tau, -1.0f);
                            sum += (double)v;
                            sumsq += (double)v * (double)v;
                            if (surf) surf[tau] = v;
                        }
                    }
                }
            }
    }
    // sum of squared deviations from the FP64 sums (as corr_mfma)
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    sumsq = block_sum<T>(sumsq, red + 64, tid);
    const double mean = sum / (double)a.L;
    const double dev = sumsq - sum * mean;
    stats_write<T>(v1, i1, v2, mean, dev > 0.0 ? dev : 0.0, a.L, tid, q.rowstats + rr.slot, 1.0f);
}

bool reduced_ucode_supported(int64_t code_len, int64_t U) {
    if (code_len <= 0 || code_len % kRedN2) return false;
    const int64_t n1 = code_len / kRedN2;
    if (n1 != 12 && n1 != 24) return false;
    return U == 2 || U == 3 || U == 4 || U == 5 || U == 8 || U == 10 || U == 12 || U == 16;
}

template <class H, int N1, int T>
static int launch_corr_upm(const CorrArgs& c, hipStream_t st) {
    constexpr size_t lds = UpmShape<N1>::lds_bytes;
    UGSDR_HIP_TRY(set_lds_limit(reinterpret_cast<const void*>(corr_upm<H, N1, T>), lds));
    hipLaunchKernelGGL((corr_upm<H, N1, T>), dim3(c.rows), dim3(T), lds, st, c);
    return UGSDR_OK;
}

template <class H, int N1>
static int launch_corr_upm_t(const CorrArgs& c, hipStream_t st, int threads) {
    return threads == 512 ? launch_corr_upm<H, N1, 512>(c, st) : launch_corr_upm<H, N1, 1024>(c, st);
}

template <class H>
static int launch_corr_upm_n(const CorrArgs& c, hipStream_t st, int threads) {
    switch (c.N / c.ucode / kRedN2) {
        case 12: return launch_corr_upm_t<H, 12>(c, st, threads);
        case 24: return launch_corr_upm_t<H, 24>(c, st, threads);
        default: set_error("reduced upsampled-code search: unsupported code length"); return UGSDR_ERR_UNSUPPORTED;
    }
}

// lengths N1 * 341 with a two-stage matrix-core plan for N1
static bool red_plan(int64_t N, int& ra, int& rb_) {
    if (N <= 0 || N % kRedN2) return false;
    switch (N / kRedN2) {
        case 60: ra = 4; rb_ = 15; return true;
        case 112: ra = 16; rb_ = 7; return true;
        case 120: ra = 8; rb_ = 15; return true;
        case 180: ra = 12; rb_ = 15; return true;
        case 240: ra = 16; rb_ = 15; return true;
        default: return false;
    }
}

// Two shapes (one workgroup fills a CU's LDS either way): 512 threads with
// two butterfly groups in flight per wave, or 1024 threads with one.
template <class H, int R1a, int R1b, int T, int GPP>
static int launch_corr_mfma(const CorrArgs& c, int grid, hipStream_t st) {
    const size_t lds = RedShape<R1a, R1b>::lds_bytes(c.N);
    UGSDR_HIP_TRY(set_lds_limit(reinterpret_cast<const void*>(corr_mfma<H, R1a, R1b, T, GPP>), lds));
    hipLaunchKernelGGL((corr_mfma<H, R1a, R1b, T, GPP>), dim3(grid), dim3(T), lds, st, c);
    return UGSDR_OK;
}

template <class H, int R1a, int R1b>
static int launch_corr_mfma_t(const CorrArgs& c, int grid, hipStream_t st, int threads) {
    return threads == 512 ? launch_corr_mfma<H, R1a, R1b, 512, 2>(c, grid, st)
                          : launch_corr_mfma<H, R1a, R1b, 1024, 1>(c, grid, st);
}

template <class H>
static int launch_corr_mfma_n(const CorrArgs& c, int grid, hipStream_t st, int threads) {
    switch (c.N / kRedN2) {
        case 60: return launch_corr_mfma_t<H, 4, 15>(c, grid, st, threads);
        case 112: return launch_corr_mfma_t<H, 16, 7>(c, grid, st, threads);
        case 120: return launch_corr_mfma_t<H, 8, 15>(c, grid, st, threads);
        case 180: return launch_corr_mfma_t<H, 12, 15>(c, grid, st, threads);
        default: return launch_corr_mfma_t<H, 16, 15>(c, grid, st, threads);
    }
}

// at most 16 workgroups of 256 threads per block (16 atomics on its word)
static dim3 absmax_grid(int64_t N, int nblk) {
    return dim3((unsigned)std::min<int64_t>((N + 4095) / 4096, 16), (unsigned)nblk);
}

int enqueue_block_max(const AcqPlan& pl, AcqSlot& s, hipStream_t st) {
    const int nblk = (int)(pl.K * pl.F);
    UGSDR_HIP_TRY(s.block_scale.reserve((size_t)nblk * 8 + (pl.ucode ? (size_t)pl.D * pl.K * 4 : 0)));
    unsigned* bmax = s.block_scale.as<unsigned>();
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, st, bmax, nblk);
    hipLaunchKernelGGL(block_absmax_kernel, absmax_grid(pl.N, nblk), dim3(256), 0, st, s.classes.as<float2>(), pl.N,
                       bmax);
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

int enqueue_reduce_classes(const AcqPlan& pl, AcqSlot& s, hipStream_t st) {
    if (pl.ucode) return UGSDR_OK;   // the upsampled-code form packs Z instead (enqueue_classes)
    const int nblk = (int)(pl.K * pl.F);
    UGSDR_HIP_TRY(s.classes_r.reserve((size_t)nblk * pl.N * 4));
    UGSDR_HIP_TRY(s.block_scale.reserve((size_t)nblk * 8));
    unsigned* bmax = s.block_scale.as<unsigned>();
    float* inv = reinterpret_cast<float*>(bmax + nblk);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(256), 0, st, bmax, nblk);
    hipLaunchKernelGGL(block_absmax_kernel, absmax_grid(pl.N, nblk), dim3(256), 0, st, s.classes.as<float2>(), pl.N,
                       bmax);
    const dim3 grid((unsigned)std::min<int64_t>((pl.N + 1023) / 1024, 64), (unsigned)nblk);
    if (pl.precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(block_pack_kernel<Fp16Ops>, grid, dim3(256), 0, st, s.classes.as<float2>(), pl.N, bmax,
                           s.classes_r.as<uint32_t>(), inv);
    else
        hipLaunchKernelGGL(block_pack_kernel<Bf16Ops>, grid, dim3(256), 0, st, s.classes.as<float2>(), pl.N, bmax,
                           s.classes_r.as<uint32_t>(), inv);
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

int enqueue_corr_reduced(const std::vector<const AcqPlan*>& plans, const std::vector<AcqSlot*>& slots,
                         hipStream_t st, DevBuf& scratch, int cus, int threads) {
    if (plans.empty() || plans.size() > (size_t)kMaxBatch) {
        set_error("a correlation batch holds 1 to " + std::to_string(kMaxBatch) + " requests");
        return UGSDR_ERR_ARG;
    }
    const AcqPlan& p0 = *plans[0];
    CorrArgs c{};
    int rows = 0;
    for (size_t i = 0; i < plans.size(); ++i) {
        const AcqPlan& pl = *plans[i];
        AcqSlot& s = *slots[i];
        if (pl.bucket() != p0.bucket() || pl.L != p0.L || pl.precision != p0.precision) {
            set_error("batch mixes transform shapes");
            return UGSDR_ERR_ARG;
        }
        char* dp = s.tables.as<char>();
        ReqDesc& q = c.reqs[i];
        q.specs = reinterpret_cast<const float2* const*>(dp + pl.off_specs);
        q.cls = reinterpret_cast<const int*>(dp + pl.off_cls);
        q.shift = reinterpret_cast<const int*>(dp + pl.off_shift);
        q.rowstats = s.rows.as<RowStat>();
        q.surfaces = s.surf.ptr ? s.surf.as<float>() : nullptr;
        q.K = (int)pl.K;
        q.F = pl.F;
        q.D = (int)pl.D;
        q.P = (int)pl.P;
        q.row_start = rows;
        if (pl.ucode) {
            // Z and its per-(bin, block) scales; the 1/U of the phase
            // transforms' normalisation (1/Lc against the search's 1/N)
            // folds into the replica scale
            q.X = reinterpret_cast<const float2*>(s.zbuf.ptr);
            q.block_inv = reinterpret_cast<const float*>(s.block_scale.as<unsigned>() + 2 * pl.K * pl.F);
            q.spec_inv = (float)(1.0 / ((double)pl.spec_scale * (double)pl.ucode));
        } else {
            q.X = reinterpret_cast<const float2*>(s.classes_r.ptr);
            q.block_inv = reinterpret_cast<const float*>(s.block_scale.as<unsigned>() + pl.K * pl.F);
            q.spec_inv = 1.0f / pl.spec_scale;
        }
        rows += (int)(pl.P * pl.D);
    }
    c.N = (int)p0.N;
    c.L = (int)p0.L;
    c.fold_mode = p0.fold_mode;
    c.rows = rows;
    c.n_req = (int)plans.size();
    c.inv_n = 1.0f;
    if (p0.ucode) {
        // one workgroup per row, Doppler-bin-major, each XCD's resident
        // workgroups (one per CU) on consecutive rows
        c.ucode = p0.ucode;
        c.row_group = cus % 8 == 0 ? cus : 0;
        const int rc = p0.precision == UGSDR_PREC_FP16 ? launch_corr_upm_n<Fp16Ops>(c, st, threads)
                                                       : launch_corr_upm_n<Bf16Ops>(c, st, threads);
        if (rc) return rc;
        UGSDR_HIP_TRY(hipGetLastError());
        return UGSDR_OK;
    }
    int max_k = 1;
    for (const AcqPlan* pl : plans) max_k = std::max(max_k, (int)pl->K);
    c.max_k = max_k;
    const int grid = std::min(c.rows, cus);
    UGSDR_HIP_TRY(scratch.reserve(reduced_scratch_bytes(p0.N, max_k, grid)));
    c.scratch = scratch.as<float2>();
    int rc = p0.precision == UGSDR_PREC_FP16 ? launch_corr_mfma_n<Fp16Ops>(c, grid, st, threads)
                                             : launch_corr_mfma_n<Bf16Ops>(c, grid, st, threads);
    if (rc) return rc;
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

size_t reduced_scratch_bytes(int64_t N, int max_k, int grid) {
    return (size_t)red_wg_words(N, max_k) * 4 * (size_t)grid;
}

bool reduced_supported(int64_t N, int64_t L) {
    int ra, rb_;
    return red_plan(N, ra, rb_) && L == N;
}

bool reduced_plan_ok(const AcqPlan& pl) { return pl.ucode ? pl.L == pl.N : reduced_supported(pl.N, pl.L); }

int code_spec_reduce(ugsdr_hip_ctx* ctx, Replica& rep, int precision) {
    // Parseval for +-1 chips: the RMS bin magnitude is sqrt(code_len)
    const float scale = (float)(1.0 / std::sqrt((double)rep.code_len));
    void* dst = nullptr;
    UGSDR_HIP_TRY(hipMalloc(&dst, (size_t)rep.code_len * 4));
    const int blocks = (int)std::min<int64_t>((rep.code_len + 255) / 256, 4096);
    if (precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(to_half2_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.code_spec), static_cast<__half2*>(dst), rep.code_len, scale);
    else
        hipLaunchKernelGGL(to_bf162_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.code_spec), static_cast<__hip_bfloat162*>(dst),
                           rep.code_len, scale);
    const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(ctx->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        (void)hipFree(dst);
        set_error(std::string("code spectrum packing: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        return UGSDR_ERR_HIP;
    }
    // the phase transforms read it in the prime-factor storage order
    if (int rc = upm_permute(static_cast<uint32_t*>(dst), 1, rep.code_len, ctx->stream)) {
        (void)hipFree(dst);
        return rc;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipFree(dst);
        set_error("code spectrum permutation failed");
        return UGSDR_ERR_HIP;
    }
    UGSDR_HIP_TRY(hipFree(rep.code_spec));
    rep.code_spec = dst;
    rep.code_scale = scale;
    return UGSDR_OK;
}

int acquire_reduced(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, ugsdr_acq_result* out, float* surfaces,
                    float* peaks) {
    UGSDR_CHECK_ARG(r->precision == UGSDR_PREC_FP16 || r->precision == UGSDR_PREC_BF16, "unknown precision");
    AcqPlan pl;
    if (int rc = make_acq_plan(ctx, r, pl)) return rc;
    if (!reduced_plan_ok(pl)) {
        set_error("reduced precision runs the four-step lengths N1 * 341 (N1 in 60, 112, 120, 180, 240), or "
                  "replicas of one code period of N1 * 341 chips (N1 in 12, 24) upsampled by an integer factor, "
                  "with surface_len == n_fft; got n_fft " + std::to_string(pl.N) + ", surface_len " +
                  std::to_string(pl.L));
        return UGSDR_ERR_UNSUPPORTED;
    }
    ctx->graph.drop();   // this call may move buffers a cached graph refers to
    ctx->graph.pending.clear();
    ctx->last_n = pl.N;
    ctx->last_periods = (int)pl.Tsub;
    ctx->last_ucode = pl.ucode;
    AcqSlot& s = ctx->slot;
    const bool keep_surf = surfaces || peaks;
    if (!keep_surf) s.surf.release();
    hipStream_t st = ctx->stream;
    ctx->mark(0);
    if (int rc = enqueue_front(pl, r, s, st, keep_surf)) return rc;
    ctx->mark(1);
    if (int rc = enqueue_classes(pl, s, st)) return rc;
    if (int rc = enqueue_reduce_classes(pl, s, st)) return rc;
    ctx->mark(2);
    const auto tuned = ctx->tuned.find(pl.N);
    const int threads = tuned != ctx->tuned.end() ? tuned->second : 1024;
    if (int rc = enqueue_corr_reduced({&pl}, {&s}, st, ctx->scratch, device_cu_count(ctx->device), threads))
        return rc;
    ctx->mark(3);
    if (int rc = enqueue_final(pl, s, st, peaks != nullptr)) return rc;
    ctx->mark(4);
    UGSDR_HIP_TRY(hipMemcpyAsync(out, s.results.ptr, (size_t)pl.P * sizeof(ugsdr_acq_result), hipMemcpyDeviceToHost, st));
    if (peaks)
        UGSDR_HIP_TRY(hipMemcpyAsync(peaks, s.results.as<char>() + (size_t)pl.P * sizeof(ugsdr_acq_result),
                                     (size_t)pl.P * pl.L * 4, hipMemcpyDeviceToHost, st));
    if (surfaces)
        UGSDR_HIP_TRY(hipMemcpyAsync(surfaces, s.surf.ptr, (size_t)pl.P * pl.D * pl.L * 4, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    if (ctx->profile)
        for (int i = 0; i < 4; ++i) UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[i], ctx->ev[i], ctx->ev[i + 1]));
    return UGSDR_OK;
}

// ------------------------------------------------------------------ DFT operator
// Batched two-stage matrix-core DFT of n = Ra * Rb points (16-bit operands,
// FP32 accumulation): the reduced-precision DiscreteFourierTransform.
template <class H, int Ra, int Rb>
__global__ __launch_bounds__(256) void dft2_mfma_kernel(const float2* in, float2* out, int batch, int inverse,
                                                         float in_scale, float out_scale) {
    constexpr int n = Ra * Rb, T = 256, NW = T / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* bufA = reinterpret_cast<uint32_t*>(smem);
    uint32_t* bufB = bufA + kRedCap / 2;
    uint4* bta = reinterpret_cast<uint4*>(bufB + kRedCap / 2);
    uint4* btb = bta + MfmaShape<Ra>::ENTRIES;
    float2* tw = reinterpret_cast<float2*>(btb + MfmaShape<Rb>::ENTRIES);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    mfma_btab_build<Ra, H>(bta, inverse != 0, 1.f, tid, T);
    mfma_btab_build<Rb, H>(btb, inverse != 0, 1.f, tid, T);
    tw_table_build(tw, n, inverse != 0, tid, T);
    const int per_wg = (kRedCap / 2) / n;
    const int s0 = blockIdx.x * per_wg, nb = min(per_wg, batch - s0);
    for (int i = tid; i < nb * n; i += T) {
        const float2 v = in[(int64_t)s0 * n + i];
        bufA[i] = pack16<H>(make_float2(v.x * in_scale, v.y * in_scale));
    }
    __syncthreads();
    mfma_stage<Ra, H, true, NW>(bufA, n, bufB, n, n, nb, nullptr, bta, lane, wave);
    __syncthreads();
    mfma_stage<Rb, H, false, NW>(bufB, n, bufA, n, n, nb, tw, btb, lane, wave);
    __syncthreads();
    for (int i = tid; i < nb * n; i += T) {
        const float2 v = unpack16<H>(bufA[i]);
        out[(int64_t)s0 * n + i] = make_float2(v.x * out_scale, v.y * out_scale);
    }
}

template <class H, int Ra, int Rb>
static int launch_dft2(const float2* in, float2* out, int batch, int inverse, float in_scale, float out_scale,
                       hipStream_t st) {
    constexpr int n = Ra * Rb;
    const int per_wg = (kRedCap / 2) / n;
    const size_t lds = (size_t)kRedCap * 4 + (size_t)(MfmaShape<Ra>::ENTRIES + MfmaShape<Rb>::ENTRIES) * 16 +
                       (size_t)n * 8;
    UGSDR_HIP_TRY(set_lds_limit(reinterpret_cast<const void*>(dft2_mfma_kernel<H, Ra, Rb>), lds));
    hipLaunchKernelGGL((dft2_mfma_kernel<H, Ra, Rb>), dim3((batch + per_wg - 1) / per_wg), dim3(256), lds, st, in,
                       out, batch, inverse, in_scale, out_scale);
    return UGSDR_OK;
}

template <class H>
static int launch_dft2_n(int n, const float2* in, float2* out, int batch, int inverse, float is, float os,
                         hipStream_t st) {
    switch (n) {
        case 60: return launch_dft2<H, 4, 15>(in, out, batch, inverse, is, os, st);
        case 112: return launch_dft2<H, 16, 7>(in, out, batch, inverse, is, os, st);
        case 120: return launch_dft2<H, 8, 15>(in, out, batch, inverse, is, os, st);
        case 180: return launch_dft2<H, 12, 15>(in, out, batch, inverse, is, os, st);
        case 240: return launch_dft2<H, 16, 15>(in, out, batch, inverse, is, os, st);
        case 341: return launch_dft2<H, 11, 31>(in, out, batch, inverse, is, os, st);
        default:
            set_error("reduced DFT lengths are 60, 112, 120, 180, 240 and 341");
            return UGSDR_ERR_UNSUPPORTED;
    }
}

}  // namespace ugsdr_hip

extern "C" int ugsdr_hip_dft_reduced(ugsdr_hip_ctx* ctx, int precision, int inverse, const float* in, float* out,
                                     int64_t n, int64_t batch) {
    using namespace ugsdr_hip;
    if (!ctx || !in || !out || n <= 0 || batch <= 0 || batch > (1 << 20)) {
        set_error("bad arguments");
        return UGSDR_ERR_ARG;
    }
    if (precision != UGSDR_PREC_FP16 && precision != UGSDR_PREC_BF16) {
        set_error("reduced DFT precision is FP16 or BF16");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)n * batch * 8;
    // power-of-two input scale to [0.5, 1); forward unscaled, inverse 1/n
    float mx = 0.f;
    for (int64_t i = 0; i < 2 * n * batch; ++i) mx = std::max(mx, std::fabs(in[i]));
    int e = 0;
    if (mx > 0.f) std::frexp(mx, &e);
    const float is = std::ldexp(1.f, -e);
    const float os = std::ldexp(1.f, e) / (inverse ? (float)n : 1.f);
    UGSDR_HIP_TRY(ctx->aux.reserve(bytes));
    UGSDR_HIP_TRY(ctx->aux2.reserve(bytes));
    UGSDR_HIP_TRY(hipMemcpyAsync(ctx->aux.ptr, in, bytes, hipMemcpyHostToDevice, ctx->stream));
    int rc = precision == UGSDR_PREC_FP16
                 ? launch_dft2_n<ugsdr_dev::Fp16Ops>((int)n, ctx->aux.as<float2>(), ctx->aux2.as<float2>(),
                                                     (int)batch, inverse, is, os, ctx->stream)
                 : launch_dft2_n<ugsdr_dev::Bf16Ops>((int)n, ctx->aux.as<float2>(), ctx->aux2.as<float2>(),
                                                     (int)batch, inverse, is, os, ctx->stream);
    if (rc) return rc;
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, ctx->aux2.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}
