// Small DFTs on the gfx950 matrix cores for the reduced-precision path.
//
// A radix-R DFT (R <= 32) of 16 butterflies is a complex matrix product
// Y[16 x R] = X[16 x R] W[R x R], run as real v_mfma_f32_16x16x32 products
// with 16-bit operands and FP32 accumulation.  The K dimension interleaves
// real and imaginary parts, K = (re x0, im x0, re x1, im x1, ...), so each
// lane's A operand is four consecutive complex inputs, and two B tiles give
// the real and imaginary outputs:
//   B_re[(m, re)][q] =  Re W[m][q]   B_re[(m, im)][q] = -Im W[m][q]
//   B_im[(m, re)][q] =  Im W[m][q]   B_im[(m, im)][q] =  Re W[m][q]
// R <= 16 is one K step (16 complex inputs) and one 16-wide output tile
// pair; 16 < R <= 32 is two K steps and two output tile pairs.  Inputs and
// outputs past R are zero padding.
//
// Operand layout of v_mfma_f32_16x16x32_{f16,bf16} for lane l (g = l / 16):
//   A: row l % 16, K elements 8 g .. 8 g + 7
//   B: column l % 16, K elements 8 g .. 8 g + 7
//   D: column l % 16, rows 4 g .. 4 g + 3 (four FP32 values)
//
// Data between stages lives in LDS as 32-bit (re, im) pairs of 16-bit values.
// Stages are out of place (two LDS buffers), so no stage holds more than one
// group's registers at a time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_fft.hpp"

namespace ugsdr_dev {

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Fp16Ops {
    typedef _Float16 e_t;
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    typedef _Float16 v2 __attribute__((ext_vector_type(2)));
    __device__ static __forceinline__ f32x4 mfma(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
struct Bf16Ops {
    typedef __bf16 e_t;
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    typedef __bf16 v2 __attribute__((ext_vector_type(2)));
    __device__ static __forceinline__ f32x4 mfma(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <class H>
__device__ __forceinline__ uint32_t pack16(float2 v) {
    typename H::v2 h = {(typename H::e_t)v.x, (typename H::e_t)v.y};
    return __builtin_bit_cast(uint32_t, h);
}
template <class H>
__device__ __forceinline__ float2 unpack16(uint32_t u) {
    const typename H::v2 h = __builtin_bit_cast(typename H::v2, u);
    return make_float2((float)h[0], (float)h[1]);
}

template <int R>
struct MfmaShape {
    static_assert(R >= 2 && R <= 32, "matrix-core DFT radix out of range");
    static constexpr int KS = R <= 16 ? 1 : 2;  // K steps of 16 complex inputs
    static constexpr int QT = KS;               // output tile pairs of 16
    static constexpr int ENTRIES = KS * QT * 2 * 64;  // uint4 per B table
};

// B operands of a radix-R DFT for every lane: exp(sgn 2 pi i m q / R) * scale
// with sgn = +1 for the inverse.  tab[((ks * QT + qt) * 2 + tile) * 64 + lane].
template <int R, class H>
__device__ void mfma_btab_build(uint4* tab, bool inv, float scale, int tid, int nthreads) {
    using S = MfmaShape<R>;
    for (int idx = tid; idx < S::ENTRIES; idx += nthreads) {
        const int lane = idx & 63, tile = (idx >> 6) & 1, combo = idx >> 7;
        const int ks = combo / S::QT, qt = combo % S::QT;
        const int g = lane >> 4, q = qt * 16 + (lane & 15);
        typename H::v8 h;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int m = ks * 16 + 4 * g + e;
            float wr = 0.f, wi = 0.f;
            if (m < R && q < R) {
                double s, c;
                sincospi(2.0 * (double)((m * q) % R) / (double)R, &s, &c);
                wr = (float)(c * scale);
                wi = (float)((inv ? s : -s) * scale);
            }
            h[2 * e] = (typename H::e_t)(tile == 0 ? wr : wi);
            h[2 * e + 1] = (typename H::e_t)(tile == 0 ? -wi : wr);
        }
        tab[idx] = __builtin_bit_cast(uint4, h);
    }
}

// One Stockham stage of a two-stage sub-transform n = Ra * Rb over nb
// sequences (sequence s at in + s * ld_in).  FIRST (radix Ra, Ns = 1):
// butterfly jj < per = n / R reads in[jj + m per], writes out[jj R + q].
// Second (radix Rb, Ns = per): twiddle tw[m jj] (a length-n table), writes
// out[jj + q per], natural order.  The caller synchronises before and after.
template <int R, class H, bool FIRST, int NW, int GPP = 2>
__device__ __forceinline__ void mfma_stage(const uint32_t* in, int ld_in, uint32_t* out, int ld_out, int n, int nb,
                                           const float2* tw, const uint4* btab, int lane, int wave) {
    using S = MfmaShape<R>;
    constexpr int KS = S::KS, QT = S::QT;
    const int per = n / R, Bt = nb * per;
    const int i = lane & 15, g = lane >> 4;
    typename H::v8 Bre[KS][QT], Bim[KS][QT];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            Bre[ks][qt] = __builtin_bit_cast(typename H::v8, btab[((ks * QT + qt) * 2 + 0) * 64 + lane]);
            Bim[ks][qt] = __builtin_bit_cast(typename H::v8, btab[((ks * QT + qt) * 2 + 1) * 64 + lane]);
        }
    // GPP groups per pass (grp, grp + NW, ...): their LDS loads and matrix
    // products overlap instead of running back to back
    for (int grp = wave; grp * 16 < Bt; grp += GPP * NW) {
        typename H::v8 A[GPP][KS];
#pragma unroll
        for (int gi = 0; gi < GPP; ++gi) {
            const int j = (grp + gi * NW) * 16 + i;
            const int jc = j < Bt ? j : 0;
            const int seq = jc / per, jj = jc - seq * per;
            const uint32_t* src = in + seq * ld_in + jj;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = ks * 16 + 4 * g + e;
                    float2 x = make_float2(0.f, 0.f);
                    if (m < R) {
                        x = unpack16<H>(src[m * per]);
                        if (!FIRST && m > 0) x = cmul(x, tw[m * jj]);
                    }
                    A[gi][ks][2 * e] = (typename H::e_t)x.x;
                    A[gi][ks][2 * e + 1] = (typename H::e_t)x.y;
                }
            }
        }
        f32x4 Dre[GPP][QT], Dim[GPP][QT];
#pragma unroll
        for (int gi = 0; gi < GPP; ++gi)
#pragma unroll
            for (int qt = 0; qt < QT; ++qt) {
                const f32x4 z = {0.f, 0.f, 0.f, 0.f};
                Dre[gi][qt] = H::mfma(A[gi][0], Bre[0][qt], z);
                Dim[gi][qt] = H::mfma(A[gi][0], Bim[0][qt], z);
#pragma unroll
                for (int ks = 1; ks < KS; ++ks) {
                    Dre[gi][qt] = H::mfma(A[gi][ks], Bre[ks][qt], Dre[gi][qt]);
                    Dim[gi][qt] = H::mfma(A[gi][ks], Bim[ks][qt], Dim[gi][qt]);
                }
            }
#pragma unroll
        for (int gi = 0; gi < GPP; ++gi)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int jo = (grp + gi * NW) * 16 + 4 * g + r;
                if (jo < Bt) {
                    const int so = jo / per, jjo = jo - so * per;
                    uint32_t* dst = out + so * ld_out + (FIRST ? jjo * R : jjo);
#pragma unroll
                    for (int qt = 0; qt < QT; ++qt) {
                        const int q = qt * 16 + i;
                        if (q < R) dst[FIRST ? q : q * per] = pack16<H>(make_float2(Dre[gi][qt][r], Dim[gi][qt][r]));
                    }
                }
            }
    }
}

// exp(sgn 2 pi i j / n) for j < n (FP64-built)
__device__ __forceinline__ void tw_table_build(float2* tw, int n, bool inv, int tid, int nthreads) {
    for (int j = tid; j < n; j += nthreads) {
        double s, c;
        sincospi(2.0 * (double)j / (double)n, &s, &c);
        tw[j] = make_float2((float)c, (float)(inv ? s : -s));
    }
}

}  // namespace ugsdr_dev
