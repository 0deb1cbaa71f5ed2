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

// x * c for two 16-bit complex pairs, rounded once more to 16 bits.  FP16
// runs on the packed half-precision pipe (v_pk_mul_f16 + v_pk_fma_f16 with
// swizzled operands); BF16 widens to FP32 (cheap: a bf16 is the top half of
// a float) and narrows back.
template <class H>
__device__ __forceinline__ uint32_t cmul16(uint32_t a, uint32_t b) {
    return pack16<H>(cmul(unpack16<H>(a), unpack16<H>(b)));
}
template <>
__device__ __forceinline__ uint32_t cmul16<Fp16Ops>(uint32_t a, uint32_t b) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 x = __builtin_bit_cast(h2, a), c = __builtin_bit_cast(h2, b);
    const h2 p = x * c.xx;                  // (xr cr, xi cr)
    const h2 cs = {-c.y, c.y};              // (-ci, ci)
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_fma(x.yx, cs, p));
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

// Start of sequence s in a bank-skewed LDS buffer with row stride LD
// (LD % 32 == 1).  LDS serves a wave's 32-bit accesses per half-wave from
// 32 banks; the skew (12 words when bit 2 of s is set, 4 more for bit 4)
// makes the three access shapes of the matrix-core stages conflict free:
// a tile's operand reads (16 sequences x lane groups m, m + 4 at an odd
// element stride), its output writes (16 consecutive outputs x sequences
// s, s + 4) and copies that walk 32 consecutive sequences.
template <int LD>
struct SeqLayout {
    static_assert(LD % 32 == 1, "row stride must be 1 mod 32");
    static constexpr int SKEW = 16;
    __host__ __device__ static constexpr int skew(int s) { return 12 * ((s >> 2) & 1) + 4 * ((s >> 4) & 1); }
    __host__ __device__ static constexpr int base(int s) { return s * LD + skew(s); }
};

// One matrix-core Stockham stage of the correlation path over nb sequences
// laid out by Lay.  Butterfly jj < PER of a sequence takes its R inputs at
// Lay::base(seq) + jj + m * MS, m < R (MS odd), as stored: 16-bit pairs with
// no conversion and no twiddle (the previous stage applied it to its FP32
// outputs).  A tile is one butterfly jj of 16 sequences sb * 16 + row: row
// i of the A operand is sequence sb * 16 + i, and lane (i, g = lane / 16)
// holds outputs q = qt * 16 + i of sequences sb * 16 + 4 g + r, r < 4 (for
// rows past nb the values are garbage and must not be stored).  Every tile
// goes to epi(pass, gi, sb, jj, Dre, Dim) with Dre / Dim the f32x4[QT]
// real / imaginary outputs; (pass, gi) name the lane's register slot
// (MAXP > 0 unrolls at most MAXP passes so slots are compile-time; 0 loops
// at run time with pass always 0).  Tile t of a wave: wave + (pass GPP + gi)
// NW, sb = t / PER, jj = t % PER.
template <int R, class H, int NW, int GPP, int MAXP, int PER, int MS, class Lay, class Epi>
__device__ __forceinline__ void mfma_tiles(const uint32_t* in, int nb, const uint4* btab, int lane, int wave,
                                           Epi epi) {
    using S = MfmaShape<R>;
    constexpr int KS = S::KS, QT = S::QT;
    const int ntile = ((nb + 15) >> 4) * PER;
    const int i = lane & 15, g = lane >> 4;
    typename H::v8 Bre[KS][QT], Bim[KS][QT];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            Bre[ks][qt] = __builtin_bit_cast(typename H::v8, btab[((ks * QT + qt) * 2 + 0) * 64 + lane]);
            Bim[ks][qt] = __builtin_bit_cast(typename H::v8, btab[((ks * QT + qt) * 2 + 1) * 64 + lane]);
        }
    auto one_pass = [&](int pass, int t0) {
        typename H::v8 A[GPP][KS];
        int sbs[GPP], jjs[GPP];
#pragma unroll
        for (int gi = 0; gi < GPP; ++gi) {
            const int t = t0 + gi * NW;
            const int tc = t < ntile ? t : 0;
            const int sb = tc / PER, jj = tc - sb * PER;
            sbs[gi] = sb;
            jjs[gi] = jj;
            int seq = sb * 16 + i;
            seq = seq < nb ? seq : sb * 16;
            const uint32_t* src = in + Lay::base(seq) + jj + 4 * g * MS;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                uint4 v;
                const int m0 = ks * 16 + 4 * g;
                v.x = m0 + 0 < R ? src[(ks * 16 + 0) * MS] : 0u;
                v.y = m0 + 1 < R ? src[(ks * 16 + 1) * MS] : 0u;
                v.z = m0 + 2 < R ? src[(ks * 16 + 2) * MS] : 0u;
                v.w = m0 + 3 < R ? src[(ks * 16 + 3) * MS] : 0u;
                A[gi][ks] = __builtin_bit_cast(typename H::v8, v);
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
            if (t0 + gi * NW < ntile) epi(pass, gi, sbs[gi], jjs[gi], Dre[gi], Dim[gi]);
    };
    if constexpr (MAXP > 0) {
#pragma unroll
        for (int pass = 0; pass < MAXP; ++pass) {
            const int t0 = wave + pass * GPP * NW;
            if (t0 < ntile) one_pass(pass, t0);
        }
    } else {
        for (int t0 = wave; t0 < ntile; t0 += GPP * NW) one_pass(0, t0);
    }
}

// The tile a mfma_tiles call with MAXP > 0 left in register slot
// (pass, gi) of a wave, or -1 if that slot holds none.
template <int NW, int GPP, int PER>
__device__ __forceinline__ int mfma_slot_tile(int pass, int gi, int nb, int wave) {
    const int ntile = ((nb + 15) >> 4) * PER;
    const int t0 = wave + pass * GPP * NW, t = t0 + gi * NW;
    return t0 < ntile && t < ntile ? t : -1;
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
