// Device-side mixed-radix FFT building blocks for gfx950.
//
// A transform (or a batch of equal-length transforms) lives in LDS as float2.
// Each Stockham stage is register-staged: every thread loads all butterflies
// it owns for the stage into registers, applies twiddles and the radix-R DFT,
// then (after a workgroup barrier) writes them back in place.  This keeps one
// LDS buffer per transform (N * 8 bytes), so N up to 16384 fits one CU.
//
// Twiddles exp(-2 pi i m / n) come from a two-level table in LDS
// (hi[m >> 7] * lo[m & 127]) built in FP64 per workgroup, accurate to ~1 ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ugsdr_dev {

constexpr int kLdsCap = 16384;  // complex samples held in LDS per workgroup
constexpr int kTwLo = 128;      // low-level twiddle table length

// ------------------------------------------------------------ constexpr trig
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_small(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_small(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// cos / sin of 2 pi num / den
constexpr double cx_cos_turn(long num, long den) {
    num = ((num % den) + den) % den;
    double x = 2.0 * kPi * (double)num / (double)den;
    if (x > kPi) x -= 2.0 * kPi;
    return cx_cos_small(x);
}
constexpr double cx_sin_turn(long num, long den) {
    num = ((num % den) + den) % den;
    double x = 2.0 * kPi * (double)num / (double)den;
    if (x > kPi) x -= 2.0 * kPi;
    return cx_sin_small(x);
}

// ------------------------------------------------------------ complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) / +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// multiply by the constant exp(-+ 2 pi i num/den) (sign by INV)
template <bool INV>
__device__ __forceinline__ float2 cmul_const(float2 a, float c, float s) {
    // forward: (c - i s); inverse: (c + i s)
    return INV ? make_float2(fmaf(a.x, c, -a.y * s), fmaf(a.x, s, a.y * c))
               : make_float2(fmaf(a.x, c, a.y * s), fmaf(a.y, c, -a.x * s));
}

// ------------------------------------------------------------ radix DFTs
// In place on v[0..R), natural order in and out.  Forward uses exp(-2 pi i).
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<3, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float s3 = 0.86602540378443864676f;
        float2 t = cadd(v[1], v[2]);
        float2 m = make_float2(fmaf(-0.5f, t.x, v[0].x), fmaf(-0.5f, t.y, v[0].y));
        float2 d = cscale(csub(v[1], v[2]), s3);
        float2 r = rot90<INV>(d);
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <bool INV>
struct Dft<4, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = rot90<INV>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <bool INV>
struct Dft<8, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        constexpr float h = 0.70710678118654752440f;
        o[1] = cmul_const<INV>(o[1], h, h);
        o[2] = rot90<INV>(o[2]);
        o[3] = cmul_const<INV>(o[3], -h, h);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <bool INV>
struct Dft<16, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
        Dft<8, INV>::run(e);
        Dft<8, INV>::run(o);
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            const float c = (float)cx_cos_turn(k, 16), s = (float)cx_sin_turn(k, 16);
            o[k] = cmul_const<INV>(o[k], c, s);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 8] = csub(e[k], o[k]);
        }
    }
};

// odd prime radix: symmetric form
//   X[q]   = x0 + sum_j a_j cos(2 pi j q/R) -+ i sum_j b_j sin(2 pi j q/R)
//   X[R-q] = x0 + sum_j a_j cos(...)        +- i sum_j b_j sin(...)
// a_j = x_j + x_{R-j}, b_j = x_j - x_{R-j}, j,q = 1..(R-1)/2.
template <int R>
struct OddTab {
    float c[(R - 1) / 2][(R - 1) / 2];
    float s[(R - 1) / 2][(R - 1) / 2];
};
template <int R>
constexpr OddTab<R> make_odd_tab() {
    OddTab<R> t{};
    constexpr int H = (R - 1) / 2;
    for (int q = 1; q <= H; ++q)
        for (int j = 1; j <= H; ++j) {
            t.c[q - 1][j - 1] = (float)cx_cos_turn((long)j * q, R);
            t.s[q - 1][j - 1] = (float)cx_sin_turn((long)j * q, R);
        }
    return t;
}

template <int R, bool INV>
struct DftOdd {
    static constexpr int H = (R - 1) / 2;
    // before the stage barrier: v[j] <- x_j + x_{R-j}, v[R-j] <- x_j - x_{R-j}
    __device__ __forceinline__ static void prep(float2* v) {
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const float2 a = cadd(v[j], v[R - j]);
            const float2 b = csub(v[j], v[R - j]);
            v[j] = a;
            v[R - j] = b;
        }
    }
    // after the barrier: produce the outputs pairwise, handing output q to
    // put(q, value) (q is a constant once the loops are unrolled)
    template <class Put>
    __device__ __forceinline__ static void store_map(const float2* v, Put put) {
        constexpr OddTab<R> tab = make_odd_tab<R>();
        float2 sum = v[0];
#pragma unroll
        for (int j = 1; j <= H; ++j) sum = cadd(sum, v[j]);
        put(0, sum);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            float2 re = v[0], im = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const float c = tab.c[q - 1][j - 1];
                const float s = tab.s[q - 1][j - 1];
                re.x = fmaf(v[j].x, c, re.x);
                re.y = fmaf(v[j].y, c, re.y);
                im.x = fmaf(v[R - j].x, s, im.x);
                im.y = fmaf(v[R - j].y, s, im.y);
            }
            const float2 r = rot90<INV>(im);  // -i*im forward, +i*im inverse
            put(q, cadd(re, r));
            put(R - q, csub(re, r));
        }
    }
    __device__ __forceinline__ static void store(const float2* v, float2* dst, int stride) {
        store_map(v, [&](int q, float2 x) { dst[q * stride] = x; });
    }
    __device__ __forceinline__ static void run(float2* v) {
        float2 out[R];
        prep(v);
        store(v, out, 1);
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = out[q];
    }
};
// Stage interface: prep() runs before the stage barrier on the loaded,
// twiddled inputs; store() runs after it and writes the R outputs.
template <int R, bool INV>
struct StageOp {
    __device__ __forceinline__ static void prep(float2* v) { Dft<R, INV>::run(v); }
    __device__ __forceinline__ static void store(const float2* v, float2* dst, int stride) {
#pragma unroll
        for (int q = 0; q < R; ++q) dst[q * stride] = v[q];
    }
};
template <int R, bool INV>
struct StageOdd {
    __device__ __forceinline__ static void prep(float2* v) { DftOdd<R, INV>::prep(v); }
    __device__ __forceinline__ static void store(const float2* v, float2* dst, int stride) {
        DftOdd<R, INV>::store(v, dst, stride);
    }
};
template <int R, bool INV> struct Stage : StageOp<R, INV> {};
template <bool INV> struct Stage<5, INV> : StageOdd<5, INV> {};
template <bool INV> struct Stage<7, INV> : StageOdd<7, INV> {};
template <bool INV> struct Stage<11, INV> : StageOdd<11, INV> {};
template <bool INV> struct Stage<13, INV> : StageOdd<13, INV> {};
template <bool INV> struct Stage<19, INV> : StageOdd<19, INV> {};
template <bool INV> struct Stage<31, INV> : StageOdd<31, INV> {};

template <bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};
template <bool INV> struct Dft<11, INV> : DftOdd<11, INV> {};
template <bool INV> struct Dft<13, INV> : DftOdd<13, INV> {};
template <bool INV> struct Dft<19, INV> : DftOdd<19, INV> {};
template <bool INV> struct Dft<31, INV> : DftOdd<31, INV> {};

// 33-point DFT as a 3 x 11 prime-factor (Good-Thomas) transform, which needs
// no twiddles between its two levels.  Input n = (11 n1 + 3 n2) mod 33 feeds
// DFT-3 number n2 at position n1; output (k1, k2) of the DFT-11s is X[k] with
// k = (22 k1 + 12 k2) mod 33 (the CRT map: 22 = 1 mod 3 = 0 mod 11,
// 12 = 0 mod 3 = 1 mod 11).
template <bool INV>
struct Pfa33 {
    // before the stage barrier: the DFT-3s, then the DFT-11 pre-combination;
    // v[11 k1 + n2] then holds group k1
    __device__ __forceinline__ static void prep(float2* v) {
        float2 y[33];
#pragma unroll
        for (int n2 = 0; n2 < 11; ++n2) {
            float2 t[3] = {v[(3 * n2) % 33], v[(11 + 3 * n2) % 33], v[(22 + 3 * n2) % 33]};
            Dft<3, INV>::run(t);
            y[n2] = t[0];
            y[11 + n2] = t[1];
            y[22 + n2] = t[2];
        }
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) DftOdd<11, INV>::prep(y + 11 * k1);
#pragma unroll
        for (int i = 0; i < 33; ++i) v[i] = y[i];
    }
    // after the barrier: output k goes to put(k, value)
    template <class Put>
    __device__ __forceinline__ static void store(const float2* v, Put put) {
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1)
            DftOdd<11, INV>::store_map(v + 11 * k1, [&](int k2, float2 x) { put((22 * k1 + 12 * k2) % 33, x); });
    }
};

// 12-point DFT as a 4 x 3 prime-factor transform: input n = (3 n1 + 4 n2) mod 12
// feeds DFT-4 number n2 at position n1; output (k1, k2) is X[(9 k1 + 4 k2) mod 12]
// (9 = 1 mod 4 = 0 mod 3, 4 = 0 mod 4 = 1 mod 3).
template <bool INV>
struct Dft<12, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 y[4][3];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) {
            float2 t[4];
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) t[n1] = v[(3 * n1 + 4 * n2) % 12];
            Dft<4, INV>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) y[k1][n2] = t[k1];
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            Dft<3, INV>::run(y[k1]);
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) v[(9 * k1 + 4 * k2) % 12] = y[k1][k2];
        }
    }
};

// 10-point DFT as a 2 x 5 prime-factor transform: input n = (5 n1 + 2 n2) mod 10
// feeds DFT-2 number n2 at position n1; output (k1, k2) is X[(5 k1 + 6 k2) mod 10]
// (5 = 1 mod 2 = 0 mod 5, 6 = 0 mod 2 = 1 mod 5).
template <bool INV>
struct Dft<10, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 y[2][5];
#pragma unroll
        for (int n2 = 0; n2 < 5; ++n2) {
            float2 t[2] = {v[(2 * n2) % 10], v[(5 + 2 * n2) % 10]};
            Dft<2, INV>::run(t);
            y[0][n2] = t[0];
            y[1][n2] = t[1];
        }
#pragma unroll
        for (int k1 = 0; k1 < 2; ++k1) {
            Dft<5, INV>::run(y[k1]);
#pragma unroll
            for (int k2 = 0; k2 < 5; ++k2) v[(5 * k1 + 6 * k2) % 10] = y[k1][k2];
        }
    }
};

// ------------------------------------------------------------ twiddle table
// lo[i] = exp(-2 pi i i / n), i < 128;  hi[i] = exp(-2 pi i (128 i) / n)
__device__ __forceinline__ void tw_build(float2* lo, float2* hi, int n, int tid, int nthreads) {
    const int nhi = (n + kTwLo - 1) / kTwLo;
    for (int i = tid; i < kTwLo + nhi; i += nthreads) {
        const long m = i < kTwLo ? i : (long)(i - kTwLo) * kTwLo;
        const double a = -2.0 * kPi * (double)(m % n) / (double)n;
        double s, c;
        sincos(a, &s, &c);
        if (i < kTwLo) lo[i] = make_float2((float)c, (float)s);
        else hi[i - kTwLo] = make_float2((float)c, (float)s);
    }
}

template <bool INV>
__device__ __forceinline__ float2 tw_get(const float2* lo, const float2* hi, int m) {
    const float2 w = cmul(hi[m >> 7], lo[m & (kTwLo - 1)]);
    return INV ? make_float2(w.x, -w.y) : w;
}

// ------------------------------------------------------------ stages
constexpr int rounds_for(int R, int T, int CAP) { return (CAP / R + T - 1) / T; }

// One register-staged Stockham stage over nb sequences of length n stored
// back to back in buf (nb * n <= CAP).  Ns = product of the radices already
// applied.  Inputs are loaded, twiddled and pre-combined before the barrier;
// outputs are formed and stored after it.
template <int R, int T, int CAP, bool INV, bool FIRST, int NC = 0, int NSC = 0>
__device__ __forceinline__ void fft_stage(float2* buf, int n_rt, int nb, int Ns_rt, const float2* lo,
                                          const float2* hi, int tid, int ld = 0) {
    constexpr int RM = rounds_for(R, T, CAP);
    const int n = NC > 0 ? NC : n_rt;
    const int lds = ld > 0 ? ld : n;  // sequence stride in buf
    const int Ns = NSC > 0 ? NSC : Ns_rt;
    const int per = n / R;
    const int B = nb * per;
    const int tws = n / (Ns * R);
    float2 v[RM][R];
    int obase[RM];
#pragma unroll
    for (int rr = 0; rr < RM; ++rr) {
        const int b = tid + rr * T;
        obase[rr] = -1;
        if (b < B) {
            const int seq = b / per;
            const int j = b - seq * per;
            const int k = FIRST ? 0 : j % Ns;
            const float2* src = buf + seq * lds + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[rr][r] = src[r * per];
            if (!FIRST && k) {
                const int step = k * tws;
#pragma unroll
                for (int r = 1; r < R; ++r) v[rr][r] = cmul(v[rr][r], tw_get<INV>(lo, hi, r * step));
            }
            Stage<R, INV>::prep(v[rr]);
            obase[rr] = seq * lds + (j - k) * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < RM; ++rr)
        if (obase[rr] >= 0) Stage<R, INV>::store(v[rr], buf + obase[rr], Ns);
    __syncthreads();
}

// A radix plan: up to 8 stages, radices from {2,3,4,5,7,8,11,13,16,31}.
struct RadixPlan {
    int n;
    int nst;
    int radix[8];
};

// Full transform(s) of the nb sequences in buf.  The caller has synchronised
// after filling buf; the result is in natural order and synchronised.
template <int T, int CAP, bool INV, bool FIRST>
__device__ __forceinline__ void fft_stage_any(int R, float2* buf, int n, int nb, int Ns, const float2* lo,
                                              const float2* hi, int tid, int ld) {
    switch (R) {
        case 2: fft_stage<2, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 3: fft_stage<3, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 4: fft_stage<4, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 5: fft_stage<5, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 7: fft_stage<7, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 8: fft_stage<8, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 11: fft_stage<11, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 13: fft_stage<13, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 16: fft_stage<16, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 19: fft_stage<19, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        case 31: fft_stage<31, T, CAP, INV, FIRST>(buf, n, nb, Ns, lo, hi, tid, ld); break;
        default: break;
    }
}

// Full transform(s) of the nb sequences in buf (nb * plan.n <= CAP).  The
// caller has synchronised after filling buf; the result is in natural order
// and synchronised.  The first stage needs no twiddles.
template <int T, int CAP, bool INV>
__device__ void fft_lds(float2* buf, const RadixPlan& plan, int nb, const float2* lo, const float2* hi, int tid,
                        int ld = 0) {
    if (plan.nst == 0) return;
    fft_stage_any<T, CAP, INV, true>(plan.radix[0], buf, plan.n, nb, 1, lo, hi, tid, ld);
    int Ns = plan.radix[0];
    for (int s = 1; s < plan.nst; ++s) {
        const int R = plan.radix[s];
        fft_stage_any<T, CAP, INV, false>(R, buf, plan.n, nb, Ns, lo, hi, tid, ld);
        Ns *= R;
    }
}

// ------------------------------------------------------------ compile-time plans
// A plan whose radices are template parameters: every stage is straight-line
// code with constant strides and divisors.
template <int... Rs>
struct PlanCT {
    static constexpr int N = (Rs * ...);
    static constexpr bool kCompileTime = true;
};
struct PlanRT {
    static constexpr int N = 0;
    static constexpr bool kCompileTime = false;
};

template <int T, int CAP, bool INV, int N, int NS, int R, int... Rest>
__device__ __forceinline__ void fft_ct_stages(float2* buf, int nb, const float2* lo, const float2* hi, int tid,
                                              int ld) {
    fft_stage<R, T, CAP, INV, NS == 1, N, NS>(buf, N, nb, NS, lo, hi, tid, ld);
    if constexpr (sizeof...(Rest) > 0) fft_ct_stages<T, CAP, INV, N, NS * R, Rest...>(buf, nb, lo, hi, tid, ld);
}

template <int T, int CAP, bool INV, class P>
struct Fft;

template <int T, int CAP, bool INV, int... Rs>
struct Fft<T, CAP, INV, PlanCT<Rs...>> {
    __device__ __forceinline__ static void run(float2* buf, const RadixPlan&, int nb, const float2* lo,
                                               const float2* hi, int tid, int ld = 0) {
        fft_ct_stages<T, CAP, INV, PlanCT<Rs...>::N, 1, Rs...>(buf, nb, lo, hi, tid, ld);
    }
};
template <int T, int CAP, bool INV>
struct Fft<T, CAP, INV, PlanRT> {
    __device__ __forceinline__ static void run(float2* buf, const RadixPlan& plan, int nb, const float2* lo,
                                               const float2* hi, int tid, int ld = 0) {
        fft_lds<T, CAP, INV>(buf, plan, nb, lo, hi, tid, ld);
    }
};

// n = n1 * n2 split for transforms larger than one workgroup's LDS
struct FourStep {
    int n, n1, n2;
    RadixPlan p1, p2;
};

// Returns v through an empty asm so the compiler cannot treat values derived
// from it as loop invariant: used at the top of loops that contain a whole
// transform, where hoisting the per-stage index arithmetic out of the loop
// would exceed the register budget.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Global-memory view of a pointer that was loaded from memory (such pointers
// are generic, which makes every access a flat instruction with a 64-bit
// VGPR address); global loads take an SGPR base plus a 32-bit offset.
// (The host pass of the same source sees a plain pointer: the qualified
// vector types only exist in device code.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
using gwptr = __attribute__((address_space(1))) T*;
#else
template <class T>
using gptr = const T*;
template <class T>
using gwptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
    return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ gwptr<T> as_global_w(T* p) {
    return (gwptr<T>)p;
}

// |z| with the hardware square root (v_sqrt_f32, 1 ulp).  The correctly
// rounded sqrtf expands to a dozen instructions and its temporaries; the
// magnitude the reference takes (IPP's 11-bit Abs) is coarser than either.
__device__ __forceinline__ float cabs_hw(float2 z) { return __builtin_amdgcn_sqrtf(fmaf(z.x, z.x, z.y * z.y)); }

// Ties the values to an ordering point: later code that depends on an
// opaque() issued after this cannot be scheduled before the values exist.
__device__ __forceinline__ void order_after(float& v) { asm volatile("" : "+v"(v)); }
// Returns idx through an empty asm that also takes v: loads addressed by the
// result wait until v has been computed (bounds how many loads are in flight).
__device__ __forceinline__ int after(int idx, float& v) {
    asm volatile("" : "+v"(idx), "+v"(v));
    return idx;
}

// Walks idx = tid, tid + T, ... < nb * rows as (t, r) = (idx % nb, idx / nb)
// with one division at the start.  U elements per step: every load of a
// step is issued before the first store, so U loads per thread are in
// flight instead of one.
template <int U, class Load, class Store>
__device__ __forceinline__ void walk_cols(int tid, int T, int nb, int rows, Load load, Store store) {
    int t = tid % nb, r = tid / nb;
    const int dt = T % nb, dr = T / nb;
    while (r < rows) {
        decltype(load(0, 0)) v[U];
        int ts[U], rs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ts[u] = t;
            rs[u] = r;
            if (r < rows) v[u] = load(t, r);
            t += dt;
            r += dr;
            if (t >= nb) {
                t -= nb;
                ++r;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (rs[u] < rows) store(ts[u], rs[u], v[u]);
    }
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// (value, index) max with the smaller index winning ties
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}

}  // namespace ugsdr_dev
