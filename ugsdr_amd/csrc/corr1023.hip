// Correlation kernel for the C/A-code lengths N = 1023 * M (M = 4, 8, 12, 16:
// 1 ms at 4.092, 8.184, 12.276 and 16.368 MHz, or 4 ms at 4.092 MHz).
//
// The inverse FFT of one row block runs as three Stockham stages,
// N = 31 * 33 * M, on 512 threads (M >= 8) or N / 31 rounded up to whole
// waves (M = 4):
//   stage 1 (radix 31, no twiddles) on the matrix cores: a wave takes 16
//           butterflies at a time, each lane loads 8 of their inputs
//           x[i] = X[(i + shift) mod N] * conj(C)[i] straight from global
//           memory (the product never round-trips LDS), forms the symmetric
//           pair sums and differences, and four 16x16x16 FP32 MFMA products
//           against the cos / sin tables give the 31 outputs;
//   stage 2 (radix 33 = 3 x 11 prime factor): twiddles w_1023^(r k) come
//           from a 31 x 33 table, one LDS read each;
//   stage 3 (radix M): butterfly j owns outputs m = j + q * 1023, q < M.
//           The thread keeps the same butterflies for every block, so |.|
//           (or the complex fold) accumulates in registers and the fold to a
//           surface of L = 1023 * QL samples is q -> q mod QL, in-thread.
// One transform is two LDS round trips and four workgroup barriers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "acq.hpp"
#include "corr_common.hpp"
#include "ctx.hpp"
#include "dev_fft.hpp"
#include "dev_mfma.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

constexpr int kC1023 = 1023;

// (cos, sin)(2 pi m / 31), m < 31, the values of the FP64 functions rounded
// once to FP32: the matrix-core A operands of the radix-31 stage
__constant__ float2 kCis31[31] = {
    {1.0000000000000000e+00f, 0.0000000000000000e+00f},
    {9.7952994125249448e-01f, 2.0129852008866006e-01f},
    {9.1895781162023060e-01f, 3.9435585511331855e-01f},
    {8.2076344120727629e-01f, 5.7126821509479231e-01f},
    {6.8896691907568663e-01f, 7.2479278722911988e-01f},
    {5.2896401032696239e-01f, 8.4864425749475092e-01f},
    {3.4730525284482028e-01f, 9.3775213214708042e-01f},
    {1.5142777750457670e-01f, 9.8846832432811138e-01f},
    {-5.0649168838712642e-02f, 9.9871650717105276e-01f},
    {-2.5065253225872042e-01f, 9.6807711886620429e-01f},
    {-4.4039415155763439e-01f, 8.9780453957074158e-01f},
    {-6.1210598254766257e-01f, 7.9077573693769887e-01f},
    {-7.5875812269279086e-01f, 6.5137248272222226e-01f},
    {-8.7434661614458209e-01f, 4.8530196253108104e-01f},
    {-9.5413925640004882e-01f, 2.9936312297335804e-01f},
    {-9.9486932339189504e-01f, 1.0116832198743272e-01f},
    {-9.9486932339189516e-01f, -1.0116832198743204e-01f},
    {-9.5413925640004882e-01f, -2.9936312297335821e-01f},
    {-8.7434661614458220e-01f, -4.8530196253108082e-01f},
    {-7.5875812269279108e-01f, -6.5137248272222203e-01f},
    {-6.1210598254766269e-01f, -7.9077573693769865e-01f},
    {-4.4039415155763423e-01f, -8.9780453957074169e-01f},
    {-2.5065253225872131e-01f, -9.6807711886620407e-01f},
    {-5.0649168838713551e-02f, -9.9871650717105276e-01f},
    {1.5142777750457667e-01f, -9.8846832432811138e-01f},
    {3.4730525284482028e-01f, -9.3775213214708042e-01f},
    {5.2896401032696239e-01f, -8.4864425749475103e-01f},
    {6.8896691907568652e-01f, -7.2479278722911999e-01f},
    {8.2076344120727629e-01f, -5.7126821509479242e-01f},
    {9.1895781162023060e-01f, -3.9435585511331872e-01f},
    {9.7952994125249426e-01f, -2.0129852008866114e-01f},
};

template <int M>
struct C1023Shape {
    static constexpr int N = kC1023 * M;
    static constexpr int B1 = 33 * M;   // stage-1 butterflies
    static constexpr int B2 = 31 * M;   // stage-2 butterflies
    // threads: whole waves, at least one per stage-2 butterfly.  M >= 8
    // runs one workgroup per CU (LDS bound) as 8 waves, two per SIMD, so
    // every phase splits evenly over the four SIMDs; M = 4 runs three
    // workgroups of 3 waves per CU.
    static constexpr int T = M >= 8 ? 512 : (B1 + 63) / 64 * 64;
    static constexpr int WAVES_PER_EU = M >= 8 ? 2 : 3;
    static_assert(T >= B2, "stage 2 runs one butterfly per thread");
    // the next block's first stage-1 gathers are issued during stage 3 where
    // the register budget (two waves per SIMD) holds them
    static constexpr bool PREFETCH = M >= 8;
    static constexpr int NSLOT = (kC1023 + T - 1) / T;  // stage-3 butterflies per thread
    static constexpr int G1 = (B1 + 15) / 16;      // stage-1 groups of 16 butterflies
    static constexpr int GW = (G1 + T / 64 - 1) / (T / 64);  // groups per wave
    // LDS: data, stage-2 table [31][33], stage-3 table [1023], reduction
    // scratch, stage-1 operand table [4][64]
    static constexpr size_t lds_bytes =
        (size_t)(N + 2 * kC1023) * sizeof(float2) + 128 * sizeof(double) + 256 * sizeof(float2);
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

// LDS of the three-stage kernels: data, stage-2 table, stage-3 table,
// reduction scratch, stage-1 operand table.  Ends with a barrier: stage 1
// reads its table before the first barrier of the block loop.
template <int M>
struct C1023Lds {
    float2 *buf, *tw2, *tw3, *tw1;
    double* red;
    __device__ C1023Lds(unsigned char* smem, int tid) {
        constexpr int N = C1023Shape<M>::N, T = C1023Shape<M>::T;
        buf = reinterpret_cast<float2*>(smem);
        tw2 = buf + N;         // [k * 33 + r] = exp(+2 pi i r k / 1023)
        tw3 = tw2 + kC1023;    // [j] = exp(+2 pi i j / N)
        red = reinterpret_cast<double*>(tw3 + kC1023);
        tw1 = reinterpret_cast<float2*>(red + 128);  // [s * 64 + lane]: see stage 1
        for (int i = tid; i < kC1023; i += T) {
            const int k = i / 33, r = i - k * 33;
            double s, c;
            sincospi(2.0 * (double)((r * k) % kC1023) / (double)kC1023, &s, &c);
            tw2[i] = make_float2((float)c, (float)s);
            sincospi(2.0 * (double)i / (double)N, &s, &c);
            tw3[i] = make_float2((float)c, (float)s);
        }
        // the A operand of stage 1's step s in lane l: (cos, sin)(2 pi k p / 31)
        // for output k = l % 16 and pair p = 4 s + l / 16
        for (int i = tid; i < 256; i += T) {
            const int l = i & 63, k = l & 15, p = 4 * (i >> 6) + (l >> 4);
            tw1[i] = kCis31[(k * p) % 31];
        }
        __syncthreads();
    }
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

// a load at a 32-bit byte offset from a wave-uniform base: one VGPR offset
// on the scalar base address, no 64-bit address arithmetic per lane
__device__ __forceinline__ float2 ld_off(gptr<float2> base, uint32_t byte_off) {
    return *reinterpret_cast<gptr<float2>>(reinterpret_cast<gptr<char>>(base) + byte_off);
}

// A raw buffer resource over `bytes` bytes at p (no stride, no swizzle):
// loads through it take one 32-bit lane offset plus an immediate of up to
// 4095 bytes, with no 64-bit address arithmetic, and read zero at offsets
// past `bytes` instead of touching memory there.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buf_ld2(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_off, 0, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Blocks k0 .. k1-1 of one row into the thread's surface samples
// S[s][u] = surface[tid + s T + u 1023].
//
// PFA (M = 4, 8, 16, coprime to 1023 = 31 x 33; unshifted rows, no fold):
// the transform runs as a three-factor prime-factor (Good-Thomas) transform
// N = 31 x 33 x M with no twiddle multiplies.  The spectra are stored in
// its input order (pfam_elem), so stage-1 butterfly j = n_b M + n_c reads
// input n_a at j + n_a 33M as in the Stockham order; stage 2 and stage 3
// are plain 33- and M-point DFTs over the same LDS layout, and the thread's
// output (j, u) -- j = k_b 31 + k_a -- is sample k of the CRT map
// k = k_a mod 31, k_b mod 33, u mod M (pfa_sample).
template <int M, int QL, bool PFA = false>
__device__ __forceinline__ void corr1023_blocks(float (&S)[C1023Shape<M>::NSLOT][QL], const C1023Lds<M>& Ld,
                                                const ReqDesc& q, int cls, int sh, gptr<float2> C, int k0, int k1,
                                                int fold_mode, int tid) {
    using S_ = C1023Shape<M>;
    constexpr int N = S_::N, T = S_::T, B1 = S_::B1, B2 = S_::B2, NSLOT = S_::NSLOT;
    constexpr int FOLD = M / QL;
    float2* buf = Ld.buf;
    const float2* tw2 = Ld.tw2;
    const float2* tw3 = Ld.tw3;
    const float2* tw1 = Ld.tw1;
    constexpr int W = T / 64, G1 = S_::G1, GW = S_::GW;
    // stage-1 gathers of group gi of block kb: 8 inputs per lane (see stage 1)
    auto gather = [&](int kb, int gi, float2 (&x)[8], float2 (&c)[8]) {
        const gptr<float2> Xb = as_global(q.X + ((int64_t)kb * q.F + cls) * N);
        const int tt = opaque(tid);
        const int lane = tt & 63, col = lane & 15, kq = lane >> 4;
        const int wv = __builtin_amdgcn_readfirstlane(tt >> 6);
        const int g = min(wv + gi * W, G1 - 1);
        const int j = min(g * 16 + col, B1 - 1);
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const int pp = 4 * st + kq;
            const unsigned i1 = j + pp * B1;
            const unsigned i2 = j + (pp == 0 ? 0 : 31 - pp) * B1;
            unsigned s1 = i1, s2 = i2;
            if constexpr (!PFA) {
                // the row's circular shift (the prime-factor rows have none)
                s1 += sh;
                s2 += sh;
                s1 -= s1 >= (unsigned)N ? N : 0;
                s2 -= s2 >= (unsigned)N ? N : 0;
            }
            x[2 * st] = ld_off(Xb, s1 * 8u);
            c[2 * st] = ld_off(C, i1 * 8u);
            x[2 * st + 1] = ld_off(Xb, s2 * 8u);
            c[2 * st + 1] = ld_off(C, i2 * 8u);
        }
    };
    // with PREFETCH, group 0 of the next block is gathered during this
    // block's stage 3
    float2 xv[2][8], cv[2][8];
    if constexpr (S_::PREFETCH) {
        if (k0 < k1) gather(k0, 0, xv[0], cv[0]);
    }
    for (int k = k0; k < k1; ++k) {
        const int t = opaque(tid);
        if constexpr (!S_::PREFETCH) gather(k, 0, xv[0], cv[0]);
        // ---- stage 1: radix 31 on the matrix cores.  Group g of the wave
        // holds butterflies j = 16 g + l % 16.  With the symmetric form's
        // s_p = x_p + x_{31-p}, d_p = x_p - x_{31-p} (p = 1..15; s_0 = x_0,
        // d_0 = 0) as the B operand (pair p = 4 s + l / 16 at step s) and
        // cos / sin(2 pi k p / 31) as the A operand, four 16x16x16 products
        // give a_k = sum_p cos * s_p and b_k = sum_p sin * d_p for k < 16 in
        // lane l's registers r (k = 4 (l / 16) + r, butterfly l % 16), and
        // X_k = a_k + i b_k, X_{31-k} = a_k - i b_k.
        // The gathers of group gi + 1 are issued before group gi's products
        // (group 0's at the top of the block, or with PREFETCH during the
        // previous block's stage 3), on straight-line
        // code: a wave's surplus group slot (the group count is not a
        // multiple of the wave count) and the lanes of the last group past
        // B1 gather a clamped index, skip the products and store nothing.
        {
            const int lane = t & 63, col = lane & 15, kq = lane >> 4;
            const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
            f32x4 acc[GW][4];
#pragma unroll
            for (int gi = 0; gi < GW; ++gi) {
                if (gi + 1 < GW) gather(k, gi + 1, xv[(gi + 1) & 1], cv[(gi + 1) & 1]);
                // slots below G1 on every wave need no check
                if ((gi + 1) * W <= G1 || wv + gi * W < G1) {
                    f32x4 ar = {0.f, 0.f, 0.f, 0.f}, ai = ar, br = ar, bi = ar;
#pragma unroll
                    for (int st = 0; st < 4; ++st) {
                        const float2 x1 = cmul(xv[gi & 1][2 * st], cv[gi & 1][2 * st]);
                        float2 x2 = cmul(xv[gi & 1][2 * st + 1], cv[gi & 1][2 * st + 1]);
                        if (st == 0 && kq == 0) x2 = make_float2(0.f, 0.f);
                        const float2 sp = cadd(x1, x2), dp = csub(x1, x2);
                        const float2 cs = tw1[st * 64 + lane];
                        ar = mfma4(cs.x, sp.x, ar);
                        ai = mfma4(cs.x, sp.y, ai);
                        br = mfma4(cs.y, dp.x, br);
                        bi = mfma4(cs.y, dp.y, bi);
                    }
                    acc[gi][0] = ar;
                    acc[gi][1] = ai;
                    acc[gi][2] = br;
                    acc[gi][3] = bi;
                }
            }
            __syncthreads();
#pragma unroll
            for (int gi = 0; gi < GW; ++gi) {
                const int g = wv + gi * W;
                const int j = g * 16 + col;
                if (g < G1 && j < B1) {
                    float2* dst = buf + j * 31;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int kk = 4 * kq + r;
                        const float a_re = acc[gi][0][r], a_im = acc[gi][1][r];
                        const float b_re = acc[gi][2][r], b_im = acc[gi][3][r];
                        dst[kk] = make_float2(a_re - b_im, a_im + b_re);
                        if (kk > 0) dst[31 - kk] = make_float2(a_re + b_im, a_im - b_re);
                    }
                }
            }
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
            if constexpr (!PFA) {
                // twiddles in groups of 8 reads, each group after the last one's products
                int tb = kk * 33;
#pragma unroll
                for (int r = 1; r < 33; ++r) {
                    if (r > 1 && (r - 1) % 8 == 0) tb = after(tb, u[r - 1].x);
                    u[r] = cmul(u[r], tw2[tb + r]);
                }
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
        if constexpr (S_::PREFETCH) {
            if (k + 1 < k1) gather(k + 1, 0, xv[0], cv[0]);
        }
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            // one slot at a time: the next slot's loads wait for this one's sums
            if (s > 0) order_after(S[s - 1][0]);
            const int j = opaque(t) + s * T;
            if (j < kC1023) {
                float2 w[M];
#pragma unroll
                for (int r = 0; r < M; ++r) w[r] = buf[j + r * kC1023];
                if constexpr (!PFA) {
                    float2 tw[M];
                    twiddle_powers<M>(tw3[j], tw);
#pragma unroll
                    for (int r = 1; r < M; ++r) w[r] = cmul(w[r], tw[r]);
                }
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
    const int slot = row_ref(q, local, a.ucode != 0).slot;
    row_stats_glob<T>(S, L, red, tid, q.rowstats + slot, a.inv_n);
    if (q.surfaces) {
        float* o = q.surfaces + (int64_t)slot * L;
        for (int j = tid; j < L; j += T) o[j] = S[j] * a.inv_n;
    }
}

// One row per workgroup: every block, then the statistics.  The first
// a.n_merge workgroups instead finish the tail rows from the stream-K
// partial surfaces (launched before this kernel), so that short work runs
// beside the first round of rows rather than after the last.
template <int M, int QL>
__global__ __launch_bounds__(C1023Shape<M>::T, C1023Shape<M>::WAVES_PER_EU) void corr1023(CorrArgs a) {
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
__global__ __launch_bounds__(C1023Shape<M>::T, C1023Shape<M>::WAVES_PER_EU) void corr1023_sk(CorrArgs a) {
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

// ---------------------------------------------------------- upsampled code
// Kernels of the upsampled-code search.  A row is M independent 1023-point
// inverse transforms, one per phase p (row sample tau = M q + p): phase p's
// input is Z_p[k'] conj(C)[k'], Z stored phase-major per (Doppler bin,
// block), C the code's 1023-point conj spectrum.  Each wave owns two phases
// end to end: their 66 radix-31 stage-1 butterflies (on the matrix cores in
// five groups of 16, as in the three-stage kernel), the phases' LDS region,
// and their 62 radix-33 stage-2 butterflies, whose 33 outputs each are summed
// as |.| in registers over the blocks.  A wave reads back only what it wrote,
// so the block loop has no workgroup barrier: waves of one CU drift freely
// and hide each other's latency.
//
// 1023 = 31 x 33 with 31 and 33 coprime runs as a prime-factor (Good-Thomas)
// transform, with no twiddles between the stages: stage-1 butterfly jj takes
// inputs k' = (33 m + 31 jj) mod 1023, m = 0..30, into a 31-point transform
// (output k1), stage 2 a 33-point transform over jj (output k2), and output
// (k1, k2) is sample q = (528 k1 + 496 k2) mod 1023 (528 = 1 mod 31 = 0 mod
// 33, 496 = 0 mod 31 = 1 mod 33), since k' q = 33 m k1 + 31 jj k2 mod 1023.
// Z and the code spectrum are stored in that input order (element k' at
// pfa1023_pos(k') = 33 m + jj), so a stage-1 group's 16 butterflies read
// 16 consecutive elements.
template <int M>
struct UpShape {
    static constexpr int W = M / 2;   // waves, two phases each
    static constexpr int T = 64 * W;
    static constexpr int G = 5;       // stage-1 groups of 16 butterflies per wave (66 of 80 used)
    // LDS: the phases' data, reduction scratch, the stage-1 operand table [4][64]
    static constexpr size_t lds_bytes = (size_t)M * kC1023 * sizeof(float2) + 128 * sizeof(double) + 256 * sizeof(float2);
};

template <int M>
struct UpLds {
    float2 *buf, *tw1;
    double* red;
    __device__ UpLds(unsigned char* smem, int tid) {
        constexpr int T = UpShape<M>::T;
        buf = reinterpret_cast<float2*>(smem);
        red = reinterpret_cast<double*>(buf + M * kC1023);
        tw1 = reinterpret_cast<float2*>(red + 128);
        for (int i = tid; i < 256; i += T) {
            const int l = i & 63, k = l & 15, p = 4 * (i >> 6) + (l >> 4);
            tw1[i] = kCis31[(k * p) % 31];
        }
        __syncthreads();
    }
};

// a wave's LDS writes before its later reads of other lanes' data: a wave's
// LDS instructions complete in order, so this only keeps the compiler from
// moving the accesses across each other
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_sched_barrier(0);
}

// storage position of element k' (< 1023) in prime-factor order: 33 m + jj
// with m = 16 k' mod 31, jj = 16 k' mod 33 (k' = 33 m + 31 jj mod 1023)
__host__ __device__ __forceinline__ int pfa1023_pos(int k) { return (16 * k % 31) * 33 + 16 * k % 33; }

// row sample of stage-2 output qq of the wave-lane (wave wv, lane l < 62):
// phase 2 wv + l / 31, stage-1 output k1 = l % 31, stage-2 output qq
template <int M>
__device__ __forceinline__ int upw_tau(int wv, int l, int qq) {
    const int pp = l >= 31 ? 1 : 0, kk = l - 31 * pp;
    return M * ((528 * kk + 496 * qq) % kC1023) + 2 * wv + pp;
}

// Blocks k0 .. k1-1 of the row of Doppler bin di with code spectrum C into
// the lane's 33 surface samples (lanes < 62 of every wave).
template <int M>
__device__ __forceinline__ void upw_blocks(float (&S)[33], const UpLds<M>& Ld, const ReqDesc& q, int di,
                                           gptr<float2> C, int k0, int k1, int tid) {
    constexpr int N = kC1023 * M, G = UpShape<M>::G;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    float2* region = Ld.buf + 2 * wv * kC1023;   // the wave's two phases, [33][31] each
    // stage-1 gathers of group g of block kb: butterfly b = 16 g + col of the
    // wave (phase 2 wv + b / 33, residue b % 33), 8 inputs per lane, through
    // buffer resources over the wave's two phases of Z and over the code
    // spectrum.  The lane index is laundered per block (opaque), so the
    // per-group offsets are recomputed in the loop instead of being hoisted
    // into registers.
    const __amdgpu_buffer_rsrc_t Cr = buf_rsrc((const void*)C, kC1023 * 8u);
    auto gather = [&](int t, int kb, int g, float2 (&x)[8], float2 (&c)[8]) {
        const int col = t & 15, kq = (t >> 4) & 3;
        const __amdgpu_buffer_rsrc_t Zr =
            buf_rsrc(q.X + ((int64_t)di * q.K + kb) * N + 2 * wv * kC1023, 2 * kC1023 * 8u);
        const int b = min(g * 16 + col, 65);
        const int pp = b >= 33 ? 1 : 0, jj = b - 33 * pp;
        const uint32_t zo = (uint32_t)pp * (kC1023 * 8u);
        // inputs m = 4 st + kq and 31 - m = (19 - kq) + 4 (3 - st) of
        // butterfly jj, stored in prime-factor order (element jj + 33 m);
        // steps of 4 inputs are 1056-byte immediates.  Input 31 - m of m = 0
        // is element jj + 1023, past the code spectrum (read as zero), and
        // that product is zeroed by the stage below anyway.
        const uint32_t o1 = (uint32_t)(jj + 33 * kq) * 8u, o2 = (uint32_t)(jj + 33 * (19 - kq)) * 8u;
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            x[2 * st] = buf_ld2(Zr, zo + o1 + 1056u * st);
            c[2 * st] = buf_ld2(Cr, o1 + 1056u * st);
            x[2 * st + 1] = buf_ld2(Zr, zo + o2 + 1056u * (3 - st));
            c[2 * st + 1] = buf_ld2(Cr, o2 + 1056u * (3 - st));
        }
    };
    float2 xv[2][8], cv[2][8];
    if (k0 < k1) gather(tid, k0, 0, xv[0], cv[0]);
    for (int k = k0; k < k1; ++k) {
        const int t = tid;
        const int lane = t & 63, col = lane & 15, kq = lane >> 4;
        // ---- stage 1: radix 31 on the matrix cores (see corr1023_blocks);
        // each group's outputs are stored at once: the region's previous
        // contents were read by this wave's previous stage 2, which is done.
        // Group 0's gathers were issued during the previous block's stage 2.
#pragma unroll
        for (int gi = 0; gi < G; ++gi) {
            if (gi + 1 < G) gather(t, k, gi + 1, xv[(gi + 1) & 1], cv[(gi + 1) & 1]);
            f32x4 ar = {0.f, 0.f, 0.f, 0.f}, ai = ar, br = ar, bi = ar;
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const float2 x1 = cmul(xv[gi & 1][2 * st], cv[gi & 1][2 * st]);
                float2 x2 = cmul(xv[gi & 1][2 * st + 1], cv[gi & 1][2 * st + 1]);
                if (st == 0 && kq == 0) x2 = make_float2(0.f, 0.f);
                const float2 sp = cadd(x1, x2), dp = csub(x1, x2);
                const float2 cs = Ld.tw1[st * 64 + lane];
                ar = mfma4(cs.x, sp.x, ar);
                ai = mfma4(cs.x, sp.y, ai);
                br = mfma4(cs.y, dp.x, br);
                bi = mfma4(cs.y, dp.y, bi);
            }
            const int b = gi * 16 + col;
            if (b < 66) {
                const int pp = b >= 33 ? 1 : 0, jj = b - 33 * pp;
                float2* dst = region + pp * kC1023 + jj * 31;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int kk = 4 * kq + r;
                    dst[kk] = make_float2(ar[r] - bi[r], ai[r] + br[r]);
                    if (kk > 0) dst[31 - kk] = make_float2(ar[r] + bi[r], ai[r] - br[r]);
                }
            }
            // one group's work at a time: keeps the scheduler from hoisting
            // every group's gathers (and their registers) to the top
            __builtin_amdgcn_sched_barrier(0);
        }
        wave_lds_sync();
        // ---- stage 2: radix 33 (3 x 11) per (phase, stage-1 output kk), no
        // twiddles (prime-factor map), |.| of its outputs
        {
            const int pp = lane >= 31 ? 1 : 0, kk = min(lane - 31 * pp, 30);
            const float2* src = region + pp * kC1023 + kk;
            float2 u[33];
#pragma unroll
            for (int r = 0; r < 33; ++r) u[r] = src[r * 31];
            Pfa33<true>::prep(u);
            // the next block's first gathers fly during this block's outputs
            if (k + 1 < k1) gather(t, k + 1, 0, xv[0], cv[0]);
            if (lane < 62) Pfa33<true>::store(u, [&](int qq, float2 x) { S[qq] += cabs_hw(x); });
        }
        wave_lds_sync();
    }
}

// the lanes' 33 surface samples -> row statistics (and the surface)
template <int M>
__device__ __forceinline__ void upw_stats(const float (&S)[33], double* red, const ReqDesc& q, int slot, float inv_n,
                                          int tid) {
    constexpr int T = UpShape<M>::T, L = kC1023 * M;
    const int lane = tid & 63, wv = tid >> 6;
    const bool act = lane < 62;
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
    if (act) {
#pragma unroll
        for (int qq = 0; qq < 33; ++qq) {
            top2_merge(v1, i1, v2, S[qq], upw_tau<M>(wv, lane, qq), -1.0f);
            sum += (double)S[qq];
        }
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double qv = 0.0;
    if (act) {
#pragma unroll
        for (int qq = 0; qq < 33; ++qq) {
            const double d = (double)S[qq] - mean;
            qv += d * d;
        }
    }
    qv = block_sum<T>(qv, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, qv, L, tid, q.rowstats + slot, inv_n);
    if (q.surfaces && act) {
        float* o = q.surfaces + (int64_t)slot * L;
#pragma unroll
        for (int qq = 0; qq < 33; ++qq) o[upw_tau<M>(wv, lane, qq)] = S[qq] * inv_n;
    }
}

// One row per workgroup (Doppler-bin-major rows, XCD-grouped); the first
// a.n_merge workgroups finish the stream-K tail rows instead.
template <int M>
__global__ __launch_bounds__(UpShape<M>::T, 2) void corr1023w(CorrArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x < a.n_merge) {
        float* S = reinterpret_cast<float*>(smem);
        double* red = reinterpret_cast<double*>(S + (a.L + 1) / 2 * 2);
        corr1023_merge_row<UpShape<M>::T>(a, blockIdx.x, a.sk_grid, S, red, tid);
        return;
    }
    const UpLds<M> Ld(smem, tid);
    const int row = up_row_of(blockIdx.x - a.n_merge, a.full_rows, a.row_group);
    const ReqDesc q = a.reqs[find_req(a, row)];
    const RowRef rr = row_ref(q, row - q.row_start, true);
    float S[33];
#pragma unroll
    for (int i = 0; i < 33; ++i) S[i] = 0.f;
    upw_blocks<M>(S, Ld, q, rr.di, as_global(q.specs[rr.p]), 0, q.K, tid);
    upw_stats<M>(S, Ld.red, q, rr.slot, a.inv_n, tid);
}

// Stream-K over the tail rows, as corr1023_sk: partial surfaces in natural
// order for corr1023_merge_row.
template <int M>
__global__ __launch_bounds__(UpShape<M>::T, 2) void corr1023w_sk(CorrArgs a) {
    constexpr int L = kC1023 * M;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const UpLds<M> Ld(smem, tid);
    const int w = blockIdx.x;
    const int u_begin = sk_begin(a.units, gridDim.x, w), u_end = sk_begin(a.units, gridDim.x, w + 1);
    for (int u = u_begin; u < u_end;) {
        const ReqDesc q = a.reqs[find_req_unit(a, u)];
        const int lu = u - q.tail_unit_start;
        const int local = q.tail_row0 + lu / q.K, k0 = lu % q.K;
        const int k1 = min(q.K, k0 + (u_end - u));
        const RowRef rr = row_ref(q, local, true);
        float S[33];
#pragma unroll
        for (int i = 0; i < 33; ++i) S[i] = 0.f;
        upw_blocks<M>(S, Ld, q, rr.di, as_global(q.specs[rr.p]), k0, k1, tid);
        float* part = reinterpret_cast<float*>(a.scratch);
        const bool whole = k0 == 0 && k1 == q.K;
        float* dst = whole ? part + ((int64_t)2 * gridDim.x + (q.row_start + local - a.full_rows)) * L
                           : part + ((int64_t)w * 2 + (u == u_begin ? 0 : 1)) * L;
        if (lane < 62) {
#pragma unroll
            for (int qq = 0; qq < 33; ++qq) dst[upw_tau<M>(wv, lane, qq)] = S[qq];
        }
        u += k1 - k0;
    }
}

template <int M>
static int launch_1023w(const CorrArgs& c, hipStream_t st) {
    const size_t lds = UpShape<M>::lds_bytes;
    constexpr int T = UpShape<M>::T;
    if (c.sk_grid > 0) {
        UGSDR_HIP_TRY(lds_attr(corr1023w_sk<M>, lds));
        hipLaunchKernelGGL((corr1023w_sk<M>), dim3(c.sk_grid), dim3(T), lds, st, c);
    }
    if (c.full_rows + c.n_merge > 0) {
        UGSDR_HIP_TRY(lds_attr(corr1023w<M>, lds));
        hipLaunchKernelGGL((corr1023w<M>), dim3(c.full_rows + c.n_merge), dim3(T), lds, st, c);
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

// workgroups of the three-stage (or the upsampled-code) kernel one CU holds (LDS bound)
static int per_cu(int M, bool up) {
    size_t lds = 0;
    switch (M) {
        case 4: lds = up ? UpShape<4>::lds_bytes : C1023Shape<4>::lds_bytes; break;
        case 8: lds = up ? UpShape<8>::lds_bytes : C1023Shape<8>::lds_bytes; break;
        case 12: lds = up ? UpShape<12>::lds_bytes : C1023Shape<12>::lds_bytes; break;
        default: lds = up ? UpShape<16>::lds_bytes : C1023Shape<16>::lds_bytes; break;
    }
    return std::max(1, (int)(160 * 1024 / lds));
}

int corr1023_grid(int64_t N, int cus, bool up) { return cus * per_cu((int)(N / kC1023), up); }

int launch_corr1023(const CorrArgs& c, int cus, DevBuf& scratch, hipStream_t st) {
    const int M = c.N / kC1023, QL = c.L / kC1023;
    if (!corr1023_supports(c.N, c.L)) {
        set_error("corr1023: unsupported length pair");
        return UGSDR_ERR_UNSUPPORTED;
    }
    const int tail_rows = c.rows - c.full_rows;
    const int G_sk = tail_rows > 0 ? std::min(corr1023_grid(c.N, cus, c.ucode != 0), c.units) : 0;
    CorrArgs a = c;
    a.sk_grid = G_sk;
    a.n_merge = G_sk > 0 ? tail_rows : 0;
    const int G = corr1023_grid(c.N, cus, c.ucode != 0);
    a.row_group = c.ucode && G % 8 == 0 ? G : 0;
    if (G_sk > 0) {
        UGSDR_HIP_TRY(scratch.reserve(((size_t)G_sk * 2 + tail_rows) * c.L * 4));
        a.scratch = scratch.as<float2>();
    }
    if (c.ucode) {
        if (c.ucode != M || QL != M) {
            set_error("corr1023: upsampled-code search of another shape");
            return UGSDR_ERR_UNSUPPORTED;
        }
        switch (M) {
            case 4: return launch_1023w<4>(a, st);
            case 8: return launch_1023w<8>(a, st);
            case 12: return launch_1023w<12>(a, st);
            default: return launch_1023w<16>(a, st);
        }
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

// Upsampled-code search: codes of 1023 chips held for U = 4, 8, 12 or 16
// samples run on the wave-autonomous kernels (U phases of 1023 points);
// codes of 1023 M chips (M = 4, 8, 12, 16) held for U = 2 ... 16 samples run
// U phase transforms of 1023 M points on the three-stage engine.
bool upsampled_code_supported(int64_t code_len, int64_t U) {
    if (code_len == kC1023) return U == 4 || U == 8 || U == 12 || U == 16;
    if (code_len % kC1023 || !corr1023_threads(code_len)) return false;
    return U == 2 || U == 3 || U == 4 || U == 5 || U == 8 || U == 10 || U == 12 || U == 16;
}

// the prime-factor engine's sample for stage-3 output u of thread slot j
// (j = k_b 31 + k_a): k = k_a mod 31, k_b mod 33, u mod M, through
// k' = (528 k_a + 496 k_b) mod 1023 (528 = 1 mod 31 = 0 mod 33, 496 = 0 mod
// 31 = 1 mod 33) and k = (1024 k' + E u) mod 1023 M with 1024 = 1 mod 1023 =
// 0 mod M and E = 0 mod 1023 = 1 mod M
template <int M>
__host__ __device__ constexpr int pfa_e() {
    int e = 0;
    while ((kC1023 * e) % M != 1) ++e;
    return kC1023 * e;
}
template <int M>
__device__ __forceinline__ int pfa_sample(int j, int u) {
    static_assert(1024 % M == 0 && kC1023 % M != 0, "M must divide 1024");
    const int ka = j % 31, kb = j / 31;
    const int k1 = (528 * ka + 496 * kb) % kC1023;
    return (1024 * k1 + pfa_e<M>() * u) % (kC1023 * M);
}

// Upsampled-code search for codes of 1023 M chips held for U samples each
// (the BOC(1,1) 4 ms code: 8184 half-chips of 10 samples at 20.46 MHz): per
// row, U phase transforms of 1023 M points (phase p's input Z_p conj(C),
// Z_p of bin d and block k at ((d K + k) U + p) 1023 M), each over every
// block on the three-stage engine (its twiddle-free prime-factor form for
// M = 4, 8, 16), its |.| sums (row samples tau = U k + p, k the output's
// sample: j + 1023 u, or pfa_sample) folded into the thread's share of the
// row statistics before the next phase.  Doppler-bin-major rows,
// XCD-grouped.
template <int M>
__global__ __launch_bounds__(C1023Shape<M>::T, C1023Shape<M>::WAVES_PER_EU) void corr1023p(CorrArgs a) {
    using S_ = C1023Shape<M>;
    constexpr int NSLOT = S_::NSLOT, T = S_::T;
    // M coprime to 1023: the twiddle-free prime-factor engine
    constexpr bool PFA = M == 4 || M == 8 || M == 16;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const C1023Lds<M> Ld(smem, tid);
    const int U = a.ucode;
    const int row = up_row_of(blockIdx.x, a.rows, a.row_group);
    const ReqDesc q0 = a.reqs[find_req(a, row)];
    const RowRef rr = row_ref(q0, row - q0.row_start, true);
    ReqDesc q = q0;
    q.F = U;
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0, sumsq = 0.0;
    float* surf = q0.surfaces ? q0.surfaces + (int64_t)rr.slot * a.L : nullptr;
    for (int p = 0; p < U; ++p) {
        float S[NSLOT][M];
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
#pragma unroll
            for (int u = 0; u < M; ++u) S[s][u] = 0.f;
        corr1023_blocks<M, M, PFA>(S, Ld, q, rr.di * q0.K * U + p, 0, as_global(q0.specs[rr.p]), 0, q0.K, 0, tid);
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            const int j = tid + s * T;
            if (j < kC1023) {
#pragma unroll
                for (int u = 0; u < M; ++u) {
                    int k = j + kC1023 * u;
                    if constexpr (PFA) k = pfa_sample<M>(j, u);
                    const int tau = U * k + p;
                    const float v = S[s][u];
                    top2_merge(v1, i1, v2, v, tau, -1.0f);
                    sum += (double)v;
                    sumsq += (double)v * (double)v;
                    if (surf) surf[tau] = v * a.inv_n;
                }
            }
        }
    }
    // sum of squared deviations from the FP64 sums (a positive surface whose
    // spread is comparable to its mean)
    block_top2<T>(v1, i1, v2, Ld.red, tid);
    sum = block_sum<T>(sum, Ld.red + 64, tid);
    sumsq = block_sum<T>(sumsq, Ld.red + 64, tid);
    const double mean = sum / (double)a.L;
    const double dev = sumsq - sum * mean;
    stats_write<T>(v1, i1, v2, mean, dev > 0.0 ? dev : 0.0, a.L, tid, q0.rowstats + rr.slot, a.inv_n);
}

int launch_corr1023p(const CorrArgs& c, int cus, hipStream_t st) {
    const int M = c.ucode > 0 ? c.N / c.ucode / kC1023 : 0;
    CorrArgs a = c;
    const int G = cus * per_cu(M, false);
    a.row_group = G % 8 == 0 ? G : 0;
    a.full_rows = a.rows;
    a.units = 0;
    a.sk_grid = 0;
    a.n_merge = 0;
    switch (M) {
        case 4:
            UGSDR_HIP_TRY(lds_attr(corr1023p<4>, C1023Shape<4>::lds_bytes));
            hipLaunchKernelGGL(corr1023p<4>, dim3(a.rows), dim3(C1023Shape<4>::T), C1023Shape<4>::lds_bytes, st, a);
            break;
        case 8:
            UGSDR_HIP_TRY(lds_attr(corr1023p<8>, C1023Shape<8>::lds_bytes));
            hipLaunchKernelGGL(corr1023p<8>, dim3(a.rows), dim3(C1023Shape<8>::T), C1023Shape<8>::lds_bytes, st, a);
            break;
        case 12:
            UGSDR_HIP_TRY(lds_attr(corr1023p<12>, C1023Shape<12>::lds_bytes));
            hipLaunchKernelGGL(corr1023p<12>, dim3(a.rows), dim3(C1023Shape<12>::T), C1023Shape<12>::lds_bytes, st,
                               a);
            break;
        case 16:
            UGSDR_HIP_TRY(lds_attr(corr1023p<16>, C1023Shape<16>::lds_bytes));
            hipLaunchKernelGGL(corr1023p<16>, dim3(a.rows), dim3(C1023Shape<16>::T), C1023Shape<16>::lds_bytes, st,
                               a);
            break;
        default: set_error("upsampled-code search: unsupported phase transform"); return UGSDR_ERR_UNSUPPORTED;
    }
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

// Upsampled-code search, the PRN-independent part: for Doppler bin d and block
// k of a request, Z[(d K + k) N + p Lc + k'] = e^(j 2 pi k'p / N) sum_u
// X_c[(k' + Lc u + m) mod N] conj(H[k' + Lc u]) e^(j 2 pi up / U), with X_c
// the block's spectrum of bin d's class c and m its integer rotation.  One
// thread per k' (the U-point inverse DFT in registers); the twiddle powers
// from e^(j 2 pi k' / N) by the same product tree as the three-stage
// kernel's last stage.  PFA (Lc = 1023, one workgroup of 1024 threads per
// (bin, block)): each phase is stored in prime-factor order, permuted
// through LDS (two buffers of four phases, one barrier per four phases) so
// the global stores stay coalesced.
// H (reduced precision, not PFA): Z is stored as 16-bit pairs scaled by
// 2^-e per (bin, block), with 2^e >= the class block's max |X| times bound
// (>= max |Z|), and zinv[(d K + k)] = 2^e.
template <int U, bool PFA, class H = void>
__global__ __launch_bounds__(PFA ? 1024 : 256) void zprep_kernel(const float2* X, const float2* conjH,
                                                                 const int* cls, const int* shift, int N, int K,
                                                                 int F, int chunks, void* Zv, const unsigned* bmax,
                                                                 float bound, float* zinv) {
    static_assert(std::is_void_v<H> || !PFA, "16-bit Z is stored in natural order");
    // PFA: two buffers of PG phases (one barrier per PG phases)
    constexpr int PG = U % 4 == 0 ? 4 : 1;
    __shared__ float2 stage[PFA ? 2 * PG * kC1023 : 1];
    const int Lc = N / U;
    const int dk = blockIdx.x / chunks, chunk = blockIdx.x - dk * chunks;
    const int d = dk / K, k = dk - d * K;
    const int kp = chunk * (int)blockDim.x + threadIdx.x;
    const bool act = kp < Lc;
    if (!PFA && !act) return;
    const float2* Xb = X + ((int64_t)k * F + cls[d]) * N;
    const int m = shift[d];
    float2 v[U];
    float2 tw[U];
    if (act) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int n = kp + Lc * u;
            int src = n + m;
            src -= src >= N ? N : 0;
            v[u] = cmul(Xb[src], conjH[n]);
        }
        Dft<U, true>::run(v);
        double sn, cs;
        sincospi(2.0 * (double)kp / (double)N, &sn, &cs);
        twiddle_powers<U>(make_float2((float)cs, (float)sn), tw);
    }
    if constexpr (!std::is_void_v<H>) {
        const float mx = __uint_as_float(bmax[k * F + cls[d]]) * bound;
        int e = 0;
        if (mx > 0.f) frexpf(mx, &e);
        const float sc = ldexpf(1.f, -e);
        if (chunk == 0 && threadIdx.x == 0) zinv[dk] = ldexpf(1.f, e);
        if (act) {
            uint32_t* z = static_cast<uint32_t*>(Zv) + (int64_t)dk * N + kp;
#pragma unroll
            for (int p = 0; p < U; ++p) {
                const float2 w = cmul(v[p], tw[p]);
                z[p * Lc] = pack16<H>(make_float2(w.x * sc, w.y * sc));
            }
        }
        return;
    }
    float2* z = static_cast<float2*>(Zv) + (int64_t)dk * N + kp;
    if constexpr (PFA) {
        // written in natural order, read back at storage position kp: the
        // element stored there is k' = (33 m + 31 jj) mod 1023 (kp = 33 m +
        // jj), 31 elements apart across a row of lanes, so both LDS passes
        // are free of bank conflicts
        const int sm = kp / 33, sj = kp - 33 * sm;
        const int src = act ? (33 * sm + 31 * sj) % kC1023 : 0;
#pragma unroll
        for (int g = 0; g < U / PG; ++g) {
            float2* b = stage + (g & 1) * PG * kC1023;
            if (act) {
#pragma unroll
                for (int i = 0; i < PG; ++i) b[i * kC1023 + kp] = cmul(v[g * PG + i], tw[g * PG + i]);
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int i = 0; i < PG; ++i) z[(g * PG + i) * kC1023] = b[i * kC1023 + src];
            }
        }
    } else {
#pragma unroll
        for (int p = 0; p < U; ++p) z[p * Lc] = cmul(v[p], tw[p]);
    }
}

// Prime-factor storage order of one 1023 M-point phase (M = 4, 8, 16): the
// element at position n_a 33M + n_b M + n_c is n = (n_a 33M + n_b 31M +
// n_c 1023) mod 1023 M (n_a < 31, n_b < 33, n_c < M).
template <int M>
__host__ __device__ inline int pfam_elem(int pos) {
    constexpr int B1 = 33 * M, N = kC1023 * M;
    const int na = pos / B1, j = pos - na * B1, nb = j / M, nc = j - nb * M;
    return (na * B1 + nb * 31 * M + nc * kC1023) % N;
}

// one phase per workgroup, in place: the phase into LDS, then each position
// from its element (coalesced both ways; the LDS gathers stride 1023
// elements, conflict-free)
template <int M>
__global__ __launch_bounds__(1024) void pfam_permute_kernel(float2* data) {
    constexpr int N = kC1023 * M;
    extern __shared__ __align__(16) unsigned char smem[];
    float2* ph = reinterpret_cast<float2*>(smem);
    float2* d = data + (int64_t)blockIdx.x * N;
    for (int i = threadIdx.x; i < N; i += 1024) ph[i] = d[i];
    __syncthreads();
    for (int pos = threadIdx.x; pos < N; pos += 1024) d[pos] = ph[pfam_elem<M>(pos)];
}

template <int M>
static int launch_pfam_permute(float2* data, int64_t phases, hipStream_t st) {
    const size_t lds = (size_t)kC1023 * M * sizeof(float2);
    UGSDR_HIP_TRY(lds_attr(pfam_permute_kernel<M>, lds));
    hipLaunchKernelGGL(pfam_permute_kernel<M>, dim3((unsigned)phases), dim3(1024), lds, st, data);
    return UGSDR_OK;
}

// the prime-factor order for phases of 1023 M points where M is 4, 8 or 16
// (other lengths stay in natural order)
static int pfam_permute(float2* data, int64_t len, int64_t phases, hipStream_t st) {
    switch (len) {
        case 4 * kC1023: return launch_pfam_permute<4>(data, phases, st);
        case 8 * kC1023: return launch_pfam_permute<8>(data, phases, st);
        case 16 * kC1023: return launch_pfam_permute<16>(data, phases, st);
        default: return UGSDR_OK;
    }
}

__global__ void pfa1023_permute_kernel(const float2* in, float2* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kC1023) out[pfa1023_pos(k)] = in[k];
}

int code_spectrum_layout(ugsdr_hip_ctx* ctx, int64_t code_len, float2* spec) {
    if (code_len != kC1023) {
        if (int rc = pfam_permute(spec, code_len, 1, ctx->stream)) return rc;
        UGSDR_HIP_TRY(hipGetLastError());
        UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return UGSDR_OK;
    }
    UGSDR_HIP_TRY(ctx->aux.reserve((size_t)kC1023 * sizeof(float2)));
    UGSDR_HIP_TRY(hipMemcpyAsync(ctx->aux.ptr, spec, (size_t)kC1023 * sizeof(float2), hipMemcpyDeviceToDevice,
                                 ctx->stream));
    hipLaunchKernelGGL(pfa1023_permute_kernel, dim3((kC1023 + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->aux.as<float2>(), spec);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

template <int U, bool PFA>
static void zprep_launch(const AcqPlan& pl, AcqSlot& s, const int* cls, const int* shift, hipStream_t st) {
    constexpr int T = PFA ? 1024 : 256;
    const int chunks = (int)((pl.N / U + T - 1) / T);
    const dim3 grid((unsigned)(pl.D * pl.K * chunks));
    if (pl.precision == UGSDR_PREC_FP32) {
        hipLaunchKernelGGL((zprep_kernel<U, PFA>), grid, dim3(T), 0, st, s.classes.as<float2>(), pl.boxcar, cls,
                           shift, (int)pl.N, (int)pl.K, pl.F, chunks, s.zbuf.ptr, nullptr, 0.f, nullptr);
        return;
    }
    if constexpr (!PFA) {
        // block_scale: [K F] class-block max bits, [K F] their inverses, then zinv [D K]
        const unsigned* bmax = s.block_scale.as<unsigned>();
        float* zinv = reinterpret_cast<float*>(s.block_scale.as<unsigned>() + 2 * pl.K * pl.F);
        if (pl.precision == UGSDR_PREC_FP16)
            hipLaunchKernelGGL((zprep_kernel<U, false, Fp16Ops>), grid, dim3(T), 0, st, s.classes.as<float2>(),
                               pl.boxcar, cls, shift, (int)pl.N, (int)pl.K, pl.F, chunks, s.zbuf.ptr, bmax,
                               pl.zbound, zinv);
        else
            hipLaunchKernelGGL((zprep_kernel<U, false, Bf16Ops>), grid, dim3(T), 0, st, s.classes.as<float2>(),
                               pl.boxcar, cls, shift, (int)pl.N, (int)pl.K, pl.F, chunks, s.zbuf.ptr, bmax,
                               pl.zbound, zinv);
    }
}

int launch_zprep(const AcqPlan& pl, AcqSlot& s, hipStream_t st) {
    const bool fp32 = pl.precision == UGSDR_PREC_FP32;
    UGSDR_HIP_TRY(s.zbuf.reserve((size_t)pl.D * pl.K * pl.N * (fp32 ? 8 : 4)));
    const char* tp = s.tables.as<char>();
    const int* cls = reinterpret_cast<const int*>(tp + pl.off_cls);
    const int* shift = reinterpret_cast<const int*>(tp + pl.off_shift);
    if (!fp32 && pl.N / pl.ucode == kC1023) {
        set_error("reduced-precision upsampled-code search of a 1023-chip code");
        return UGSDR_ERR_UNSUPPORTED;
    }
    // 1023-chip codes run on the wave-autonomous kernels (prime-factor order)
    const bool pfa = pl.N / pl.ucode == kC1023;
    switch (pl.ucode) {
        case 2: zprep_launch<2, false>(pl, s, cls, shift, st); break;
        case 3: zprep_launch<3, false>(pl, s, cls, shift, st); break;
        case 5: zprep_launch<5, false>(pl, s, cls, shift, st); break;
        case 10: zprep_launch<10, false>(pl, s, cls, shift, st); break;
        case 4: pfa ? zprep_launch<4, true>(pl, s, cls, shift, st) : zprep_launch<4, false>(pl, s, cls, shift, st); break;
        case 8: pfa ? zprep_launch<8, true>(pl, s, cls, shift, st) : zprep_launch<8, false>(pl, s, cls, shift, st); break;
        case 12:
            pfa ? zprep_launch<12, true>(pl, s, cls, shift, st) : zprep_launch<12, false>(pl, s, cls, shift, st);
            break;
        case 16:
            pfa ? zprep_launch<16, true>(pl, s, cls, shift, st) : zprep_launch<16, false>(pl, s, cls, shift, st);
            break;
        default: set_error("upsampled-code search of an unsupported factor"); return UGSDR_ERR_UNSUPPORTED;
    }
    UGSDR_HIP_TRY(hipGetLastError());
    // the reduced form's phase transforms read Z in prime-factor order, and
    // so do the FP32 ones of 1023 M points with M = 4, 8, 16
    if (!fp32) return upm_permute(s.zbuf.as<uint32_t>(), pl.D * pl.K * pl.ucode, pl.N / pl.ucode, st);
    if (!pfa) return pfam_permute(s.zbuf.as<float2>(), pl.N / pl.ucode, pl.D * pl.K * pl.ucode, st);
    return UGSDR_OK;
}

}  // namespace ugsdr_hip
