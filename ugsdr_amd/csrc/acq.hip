// Fused parallel code-phase acquisition for gfx950.
//
// Data flow of one search (all device resident):
//   front   : IQ unpack (int8/int16/float/real int8) + carrier wipe-off tone
//             (FP64 phase) + rational boxcar resampling            -> sig[K*N]
//   classes : forward FFT of every block k for every fractional Doppler class
//             phi (d*N/fs = m + phi, m integer): X[k][c] = FFT(x_k e^{-2pi i phi n/N})
//             A Doppler bin d then reads X[k][c(d)] rotated by m(d), which equals
//             FFT(x_k e^{-2pi i d n / fs}) exactly in exact arithmetic.
//   corr    : per (PRN, Doppler) row, per block: Y = X_rot * conj(C_p), inverse
//             FFT in LDS, |.| (or complex fold), fold to surface_len, sum over
//             blocks in registers; then first-index max, runner-up, mean and
//             n-1 sigma by workgroup reductions.
//   final   : per PRN, strict-< selection over Doppler bins and threshold.
// N <= 16384 runs each transform inside one workgroup's LDS; larger N uses a
// four-step split N = N1*N2 through a workgroup-private global scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "plan.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

// ------------------------------------------------------------------ front end
struct FrontArgs {
    const void* iq;
    int fmt;
    int64_t n_in;
    int64_t n_out;
    double f_turns;   // mixing frequency / fs_in (cycles per sample), 0 = no tone
    double ph_turns;  // initial phase in cycles
    int apply_tone;
    int64_t up;       // lcm / fs_in
    int64_t down;     // lcm / fs_acq
    float2* out;
};

__device__ __forceinline__ float2 load_iq(const void* p, int fmt, int64_t i) {
    switch (fmt) {
        case UGSDR_FMT_CI8: {
            const char2 v = static_cast<const char2*>(p)[i];
            return make_float2((float)v.x, (float)v.y);
        }
        case UGSDR_FMT_CI16: {
            const short2 v = static_cast<const short2*>(p)[i];
            return make_float2((float)v.x, (float)v.y);
        }
        case UGSDR_FMT_CF32: return static_cast<const float2*>(p)[i];
        default: return make_float2((float)static_cast<const signed char*>(p)[i], 0.f);
    }
}

// e^{2 pi i t} for a phase t in cycles (FP64), reduced before the FP32 sincos
__device__ __forceinline__ float2 phasor_turns(double t) {
    t -= rint(t);
    float s, c;
    sincosf((float)(2.0 * kPi * t), &s, &c);
    return make_float2(c, s);
}

__global__ void front_kernel(FrontArgs a) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < a.n_out;
         o += (int64_t)gridDim.x * blockDim.x) {
        if (a.down == 1 && a.up == 1) {
            float2 v = load_iq(a.iq, a.fmt, o);
            if (a.apply_tone) v = cmul(v, phasor_turns(fma((double)o, a.f_turns, a.ph_turns)));
            a.out[o] = v;
        } else {
            double sr = 0.0, si = 0.0;
            const int64_t u0 = o * a.down;
            for (int64_t u = u0; u < u0 + a.down; ++u) {
                const int64_t i = u / a.up;
                if (i >= a.n_in) break;
                float2 v = load_iq(a.iq, a.fmt, i);
                if (a.apply_tone) v = cmul(v, phasor_turns(fma((double)i, a.f_turns, a.ph_turns)));
                sr += v.x;
                si += v.y;
            }
            a.out[o] = make_float2((float)(sr / (double)a.down), (float)(si / (double)a.down));
        }
    }
}

// ------------------------------------------------------------------ LDS layout
__host__ __device__ inline int tw_len(int n) { return kTwLo + (n + kTwLo - 1) / kTwLo; }

// ------------------------------------------------------------------ forward FFTs
// mode 0: signal block k times the class tone; mode 1: replica from chips.
struct FwdArgs {
    RadixPlan plan;      // small path
    FourStep fs;         // large path
    int N;
    int F;
    const float2* sig;           // mode 0
    const double* class_frac;    // mode 0: phi per class
    const signed char* chips;    // mode 1
    int64_t n_chips;
    int64_t len;                 // n_chips * repeats
    const float* samples;        // mode 2: real replica samples
    float2* out;                 // [grid][N]
    float2* scratch;             // large path, per workgroup N complex
    int mode;
};

__device__ __forceinline__ float2 fwd_source(const FwdArgs& a, int blk, int64_t n) {
    if (a.mode == 0) {
        const int k = blk / a.F, c = blk % a.F;
        const float2 x = a.sig[(int64_t)k * a.N + n];
        const double phi = a.class_frac[c];
        if (phi == 0.0) return x;
        return cmul(x, phasor_turns(-phi * (double)n / (double)a.N));
    }
    if (a.mode == 1) {
        const int64_t m = ((n * a.len) / a.N) % a.len;
        return make_float2((float)a.chips[m % a.n_chips], 0.f);
    }
    return make_float2(a.samples[n], 0.f);
}

// waves per SIMD the LDS footprint allows: 4 for N <= 4096, else 2
constexpr int waves_for(int CAP) { return CAP <= 4096 ? 4 : 2; }

template <int T, int CAP, class P>
__global__ __launch_bounds__(T, waves_for(CAP)) void fwd_small(FwdArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float2* buf = reinterpret_cast<float2*>(smem);
    float2* lo = buf + a.N;
    float2* hi = lo + kTwLo;
    const int tid = threadIdx.x, blk = blockIdx.x;
    tw_build(lo, hi, a.N, tid, T);
    for (int i = tid; i < a.N; i += T) buf[i] = fwd_source(a, blk, i);
    __syncthreads();
    Fft<T, CAP, false, P>::run(buf, a.plan, 1, lo, hi, tid);
    float2* o = a.out + (int64_t)blk * a.N;
    const bool conj_out = a.mode != 0;
    for (int i = tid; i < a.N; i += T) {
        float2 v = buf[i];
        if (conj_out) v.y = -v.y;
        o[i] = v;
    }
}

// four-step: X[k1 + N1 k2] = sum_n2 W_N2^{n2 k2} W_N^{n2 k1} sum_n1 x[N2 n1 + n2] W_N1^{n1 k1}
template <int T, bool INV, class P1, class P2, class Src, class Sink>
__device__ void fourstep(float2* lds, const FourStep& f, const float2* lo1, const float2* hi1, const float2* lo2,
                         const float2* hi2, const float2* loN, const float2* hiN, float2* G, Src src, Sink sink,
                         int tid0) {
    const int N1 = P1::kCompileTime ? P1::N : f.n1;
    const int N2 = P2::kCompileTime ? P2::N : f.n2;
    const int cb = max(1, kLdsCap / N1);
    for (int c0 = 0; c0 < N2; c0 += cb) {
        const int tid = opaque(tid0);
        const int nb = min(cb, N2 - c0);
        for (int idx = tid; idx < nb * N1; idx += T) {
            const int t = idx % nb, n1 = idx / nb;
            lds[t * N1 + n1] = src((int64_t)N2 * n1 + c0 + t);
        }
        __syncthreads();
        Fft<T, kLdsCap, INV, P1>::run(lds, f.p1, nb, lo1, hi1, tid);
        for (int idx = tid; idx < nb * N1; idx += T) {
            const int t = idx % nb, k1 = idx / nb;
            const int n2 = c0 + t;
            G[(int64_t)k1 * N2 + n2] = cmul(lds[t * N1 + k1], tw_get<INV>(loN, hiN, n2 * k1));
        }
        __syncthreads();
    }
    const int rb = max(1, kLdsCap / N2);
    for (int r0 = 0; r0 < N1; r0 += rb) {
        const int tid = opaque(tid0);
        const int nb = min(rb, N1 - r0);
        for (int idx = tid; idx < nb * N2; idx += T) {
            const int t = idx / N2, n2 = idx - t * N2;
            lds[idx] = G[(int64_t)(r0 + t) * N2 + n2];
        }
        __syncthreads();
        Fft<T, kLdsCap, INV, P2>::run(lds, f.p2, nb, lo2, hi2, tid);
        for (int idx = tid; idx < nb * N2; idx += T) {
            const int t = idx % nb, k2 = idx / nb;
            sink((int64_t)(r0 + t) + (int64_t)N1 * k2, lds[t * N2 + k2]);
        }
        __syncthreads();
    }
}

struct LargeLds {
    float2 *buf, *lo1, *hi1, *lo2, *hi2, *loN, *hiN;
    double* red;
};
__device__ __forceinline__ LargeLds large_lds(unsigned char* smem, const FourStep& f) {
    LargeLds L;
    L.buf = reinterpret_cast<float2*>(smem);
    L.lo1 = L.buf + kLdsCap;
    L.hi1 = L.lo1 + kTwLo;
    L.lo2 = L.lo1 + tw_len(f.n1);
    L.hi2 = L.lo2 + kTwLo;
    L.loN = L.lo2 + tw_len(f.n2);
    L.hiN = L.loN + kTwLo;
    L.red = reinterpret_cast<double*>(L.loN + tw_len(f.n));
    return L;
}
__device__ __forceinline__ void large_tables(const LargeLds& L, const FourStep& f, int tid, int T) {
    tw_build(L.lo1, L.hi1, f.n1, tid, T);
    tw_build(L.lo2, L.hi2, f.n2, tid, T);
    tw_build(L.loN, L.hiN, f.n, tid, T);
}

template <int T, class P1, class P2>
__global__ __launch_bounds__(T, 2) void fwd_large(FwdArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const LargeLds L = large_lds(smem, a.fs);
    const int tid = threadIdx.x, blk = blockIdx.x;
    large_tables(L, a.fs, tid, T);
    __syncthreads();
    float2* G = a.scratch + (int64_t)blockIdx.x * a.N;
    float2* o = a.out + (int64_t)blk * a.N;
    const bool conj_out = a.mode != 0;
    fourstep<T, false, P1, P2>(
        L.buf, a.fs, L.lo1, L.hi1, L.lo2, L.hi2, L.loN, L.hiN, G,
        [&](int64_t n) { return fwd_source(a, blk, n); },
        [&](int64_t k, float2 v) {
            if (conj_out) v.y = -v.y;
            o[k] = v;
        },
        tid);
}

// ------------------------------------------------------------------ correlation
struct CorrArgs {
    RadixPlan plan;
    FourStep fs;
    int N, K, F, L, D, P;
    int fold_mode;          // 0: |.| then fold, 1: fold complex then |.|
    int rows;
    const float2* X;        // [K][F][N]
    const float2* const* specs;  // [P] device pointers to conj spectra
    const int* cls;         // [D]
    const int* shift;       // [D]
    RowStat* rowstats;      // [P*D]
    float* surfaces;        // optional [P*D][L]
    float2* scratch;        // large path: per workgroup 2N complex + L floats
    float inv_n;
};

// block-wide reductions over T threads (T/64 waves)
template <int T>
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
    v = wave_sum_d(v);
    const int w = tid >> 6, l = tid & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < T / 64; ++i) s += red[i];
    return s;
}

// top-2 with first-index tie break: (v1,i1) the max, v2 the runner-up value
__device__ __forceinline__ void top2_merge(float& v1, int& i1, float& v2, float ov1, int oi1, float ov2) {
    if (ov1 > v1 || (ov1 == v1 && oi1 < i1)) {
        v2 = fmaxf(v1, ov2);
        v1 = ov1;
        i1 = oi1;
    } else {
        v2 = fmaxf(v2, ov1);
    }
}

template <int T>
__device__ __forceinline__ void block_top2(float& v1, int& i1, float& v2, double* red, int tid) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float a = __shfl_xor(v1, off, 64);
        const int b = __shfl_xor(i1, off, 64);
        const float c = __shfl_xor(v2, off, 64);
        top2_merge(v1, i1, v2, a, b, c);
    }
    float* rf = reinterpret_cast<float*>(red);
    int* ri = reinterpret_cast<int*>(red);
    const int w = tid >> 6, l = tid & 63;
    __syncthreads();
    if (l == 0) {
        rf[3 * w] = v1;
        ri[3 * w + 1] = i1;
        rf[3 * w + 2] = v2;
    }
    __syncthreads();
    v1 = rf[0];
    i1 = ri[1];
    v2 = rf[2];
#pragma unroll
    for (int i = 1; i < T / 64; ++i) top2_merge(v1, i1, v2, rf[3 * i], ri[3 * i + 1], rf[3 * i + 2]);
}

// Row statistics: first-index max, runner-up, mean and n-1 sigma (two-pass).
template <int T>
__device__ __forceinline__ void stats_write(float v1, int i1, float v2, double mean, double q, int L, int tid,
                                            RowStat* out, float inv_n) {
    if (tid == 0) {
        RowStat r;
        r.level = v1 * inv_n;
        r.index = i1;
        r.second = v2 > 0.f ? v2 * inv_n : 0.f;
        r.mean = L >= 2 ? (float)(mean * inv_n) : 0.f;
        r.stddev = L >= 2 ? (float)(sqrt(q / (double)(L - 1)) * inv_n) : 0.f;
        *out = r;
    }
}

// surface held in registers: S[i] is element j = tid + i*T (j < L)
template <int T, int SR>
__device__ __forceinline__ void row_stats_reg(const float (&S)[SR], int L, double* red, int tid, RowStat* out,
                                              float inv_n) {
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
        const int j = tid + i * T;
        if (j < L) {
            top2_merge(v1, i1, v2, S[i], j, -1.0f);
            sum += (double)S[i];
        }
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double q = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
        const int j = tid + i * T;
        if (j < L) {
            const double d = (double)S[i] - mean;
            q += d * d;
        }
    }
    q = block_sum<T>(q, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, q, L, tid, out, inv_n);
}

// surface in (workgroup-private) global memory
template <int T>
__device__ __forceinline__ void row_stats_glob(const float* S, int L, double* red, int tid, RowStat* out,
                                               float inv_n) {
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
    for (int j = tid; j < L; j += T) {
        const float s = S[j];
        top2_merge(v1, i1, v2, s, j, -1.0f);
        sum += (double)s;
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double q = 0.0;
    for (int j = tid; j < L; j += T) {
        const double d = (double)S[j] - mean;
        q += d * d;
    }
    q = block_sum<T>(q, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, q, L, tid, out, inv_n);
}

template <int T, int CAP, class P>
__global__ __launch_bounds__(T, waves_for(CAP)) void corr_small(CorrArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int SR = CAP / T;
    float2* buf = reinterpret_cast<float2*>(smem);
    float2* lo = buf + a.N;
    float2* hi = lo + kTwLo;
    double* red = reinterpret_cast<double*>(lo + tw_len(a.N));
    const int tid = threadIdx.x;
    tw_build(lo, hi, a.N, tid, T);
    const int N = a.N, L = a.L, R = a.N / a.L;
    for (int row = blockIdx.x; row < a.rows; row += gridDim.x) {
        const int p = row / a.D, di = row - (row / a.D) * a.D;
        const int cls = a.cls[di], sh = a.shift[di];
        const float2* __restrict__ C = a.specs[p];
        float S[SR];
#pragma unroll
        for (int i = 0; i < SR; ++i) S[i] = 0.f;
        for (int k = 0; k < a.K; ++k) {
            const int tq = opaque(tid);
            const float2* __restrict__ Xk = a.X + ((int64_t)k * a.F + cls) * N;
            for (int i = tq; i < N; i += T) {
                int src = i + sh;
                src -= src >= N ? N : 0;
                buf[i] = cmul(Xk[src], C[i]);
            }
            __syncthreads();
            Fft<T, CAP, true, P>::run(buf, a.plan, 1, lo, hi, tq);
            if (a.fold_mode == 0) {
                if (R == 1) {
#pragma unroll
                    for (int i = 0; i < SR; ++i) {
                        const int j = tq + i * T;
                        if (j < L) {
                            const float2 v = buf[j];
                            S[i] += sqrtf(fmaf(v.x, v.x, v.y * v.y));
                        }
                    }
                } else {
                    for (int j = tq; j < N; j += T) {
                        const float2 v = buf[j];
                        buf[j].x = sqrtf(fmaf(v.x, v.x, v.y * v.y));
                    }
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < SR; ++i) {
                        const int j = tq + i * T;
                        if (j < L) {
                            float acc = buf[j].x;
                            for (int r = 1; r < R; ++r) acc += buf[j + r * L].x;
                            S[i] += acc;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < SR; ++i) {
                    const int j = tq + i * T;
                    if (j < L) {
                        float2 z = buf[j];
                        for (int r = 1; r < R; ++r) z = cadd(z, buf[j + r * L]);
                        S[i] += sqrtf(fmaf(z.x, z.x, z.y * z.y));
                    }
                }
            }
            __syncthreads();
        }
        row_stats_reg<T, SR>(S, L, red, tid, a.rowstats + row, a.inv_n);
        if (a.surfaces) {
            float* o = a.surfaces + (int64_t)row * L;
#pragma unroll
            for (int i = 0; i < SR; ++i) {
                const int j = tid + i * T;
                if (j < L) o[j] = S[i] * a.inv_n;
            }
        }
        __syncthreads();
    }
}

template <int T, class P1, class P2>
__global__ __launch_bounds__(T, 2) void corr_large(CorrArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const LargeLds Ld = large_lds(smem, a.fs);
    const int tid = threadIdx.x;
    large_tables(Ld, a.fs, tid, T);
    __syncthreads();
    const int N = a.N, L = a.L, R = a.N / a.L;
    float2* G = a.scratch + (int64_t)blockIdx.x * (2 * (int64_t)N + (L + 1) / 2);
    float2* B = G + N;
    float* A = reinterpret_cast<float*>(B);
    float* S = reinterpret_cast<float*>(B + N);
    for (int row = blockIdx.x; row < a.rows; row += gridDim.x) {
        const int p = row / a.D, di = row - (row / a.D) * a.D;
        const int cls = a.cls[di], sh = a.shift[di];
        const float2* __restrict__ C = a.specs[p];
        for (int k = 0; k < a.K; ++k) {
            const int tq = opaque(tid);
            const float2* __restrict__ Xk = a.X + ((int64_t)k * a.F + cls) * N;
            const int fm = a.fold_mode;
            fourstep<T, true, P1, P2>(
                Ld.buf, a.fs, Ld.lo1, Ld.hi1, Ld.lo2, Ld.hi2, Ld.loN, Ld.hiN, G,
                [&](int64_t n) {
                    int64_t src = n + sh;
                    src -= src >= N ? N : 0;
                    return cmul(Xk[src], C[n]);
                },
                [&](int64_t kk, float2 v) {
                    if (fm == 0) A[kk] = sqrtf(fmaf(v.x, v.x, v.y * v.y));
                    else B[kk] = v;
                },
                tq);
            for (int j = tq; j < L; j += T) {
                float acc;
                if (fm == 0) {
                    acc = A[j];
                    for (int r = 1; r < R; ++r) acc += A[j + (int64_t)r * L];
                } else {
                    float2 z = B[j];
                    for (int r = 1; r < R; ++r) z = cadd(z, B[j + (int64_t)r * L]);
                    acc = sqrtf(fmaf(z.x, z.x, z.y * z.y));
                }
                S[j] = (k == 0) ? acc : S[j] + acc;
            }
            __syncthreads();
        }
        row_stats_glob<T>(S, L, Ld.red, tid, a.rowstats + row, a.inv_n);
        if (a.surfaces) {
            float* o = a.surfaces + (int64_t)row * L;
            for (int j = tid; j < L; j += T) o[j] = S[j] * a.inv_n;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ finalize
struct FinalArgs {
    const RowStat* rowstats;
    const double* doppler;
    int D, P;
    double ratio, intermediate_frequency, threshold;
    int reshape;
    ugsdr_acq_result* out;
    const float* surfaces;  // optional, to gather peaks
    float* peaks;           // optional [P][L]
    int L;
};

__global__ void finalize_kernel(FinalArgs a) {
    const int p = blockIdx.x;
    __shared__ int sel;
    if (threadIdx.x == 0) {
        ugsdr_acq_result r;
        r.level = 0.0;
        r.sigma = 1.0;
        r.mean = 0.0;
        r.stddev = 0.0;
        r.doppler = 0.0;
        r.code_offset = 0.0;
        r.runner_up_code = 0.0;
        r.index = 0;
        r.doppler_bin = -1;
        double runner = 0.0;
        for (int di = 0; di < a.D; ++di) {
            const RowStat s = a.rowstats[(int64_t)p * a.D + di];
            const double level = (double)s.level;
            if (r.level < level) {
                if (r.doppler_bin >= 0 && r.level > runner) runner = r.level;
                r.level = level;
                r.mean = s.mean;
                r.stddev = s.stddev;
                r.sigma = (double)(s.stddev + s.mean);
                r.index = s.index;
                r.code_offset = a.ratio * (double)s.index;
                r.doppler = a.doppler[di] + a.intermediate_frequency;
                r.doppler_bin = di;
                r.runner_up_code = s.second;
            } else if (level > runner) {
                runner = level;
            }
        }
        r.runner_up_doppler = runner;
        r.snr = r.level / r.sigma;
        r.detected = (!a.reshape || r.snr > a.threshold) ? 1 : 0;
        a.out[p] = r;
        sel = r.doppler_bin;
    }
    __syncthreads();
    if (a.peaks) {
        float* o = a.peaks + (int64_t)p * a.L;
        const float* src = sel >= 0 ? a.surfaces + ((int64_t)p * a.D + sel) * a.L : nullptr;
        for (int j = threadIdx.x; j < a.L; j += blockDim.x) o[j] = src ? src[j] : 0.f;
    }
}

// ------------------------------------------------------------------ host side
static size_t small_lds_bytes(int N, int T) {
    return (size_t)N * sizeof(float2) + (size_t)tw_len(N) * sizeof(float2) + 128 * sizeof(double);
}
static size_t large_lds_bytes(const FourStep& f) {
    return (size_t)kLdsCap * sizeof(float2) +
           (size_t)(tw_len(f.n1) + tw_len(f.n2) + tw_len(f.n)) * sizeof(float2) + 128 * sizeof(double);
}

template <class K>
static hipError_t set_lds(K kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

static int g_cu_count = 0;

static int cu_count(int device) {
    if (!g_cu_count) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) g_cu_count = prop.multiProcessorCount;
        if (g_cu_count <= 0) g_cu_count = 256;
    }
    return g_cu_count;
}

// forward transform(s): `blocks` transforms of length N written to out
// Compile-time plans for the lengths of the acquisition workloads; any other
// length runs the runtime-plan kernels.
using P4092 = PlanCT<31, 11, 4, 3>;
using P8184 = PlanCT<31, 11, 8, 3>;
using P12276 = PlanCT<31, 11, 4, 3, 3>;
using P16368 = PlanCT<31, 16, 11, 3>;
using P341 = PlanCT<31, 11>;
using P60 = PlanCT<5, 4, 3>;
using P112 = PlanCT<16, 7>;
using P120 = PlanCT<8, 5, 3>;
using P180 = PlanCT<5, 4, 3, 3>;
using P240 = PlanCT<16, 5, 3>;

template <int T, int CAP, class P>
static int launch_fwd_small(ugsdr_hip_ctx* ctx, FwdArgs& a, int blocks) {
    const size_t lds = small_lds_bytes(a.N, T);
    UGSDR_HIP_TRY(set_lds(fwd_small<T, CAP, P>, lds));
    hipLaunchKernelGGL((fwd_small<T, CAP, P>), dim3(blocks), dim3(T), lds, ctx->stream, a);
    return UGSDR_OK;
}

template <class P1, class P2>
static int launch_fwd_large(ugsdr_hip_ctx* ctx, FwdArgs& a, int blocks) {
    constexpr int T = 512;
    const size_t lds = large_lds_bytes(a.fs);
    UGSDR_HIP_TRY(set_lds(fwd_large<T, P1, P2>, lds));
    hipLaunchKernelGGL((fwd_large<T, P1, P2>), dim3(blocks), dim3(T), lds, ctx->stream, a);
    return UGSDR_OK;
}

// four-step split with a 341-point second factor when N/341 has a compile-time plan
static bool fourstep_341(int N, FourStep& f) {
    if (N % 341) return false;
    const int n1 = N / 341;
    if (n1 != 60 && n1 != 112 && n1 != 120 && n1 != 180 && n1 != 240) return false;
    f.n = N;
    f.n1 = n1;
    f.n2 = 341;
    make_plan(n1, f.p1);
    make_plan(341, f.p2);
    return true;
}

// forward transform(s): `blocks` transforms of length N written to out
int launch_forward(ugsdr_hip_ctx* ctx, FwdArgs& a, int blocks) {
    int rc = UGSDR_OK;
    if (a.N <= kLdsCap) {
        if (!make_plan(a.N, a.plan)) {
            set_error("unsupported transform length " + std::to_string(a.N));
            return UGSDR_ERR_UNSUPPORTED;
        }
        switch (a.N) {
            case 4092: rc = launch_fwd_small<256, 4096, P4092>(ctx, a, blocks); break;
            case 8184: rc = launch_fwd_small<256, 8192, P8184>(ctx, a, blocks); break;
            case 12276: rc = launch_fwd_small<512, kLdsCap, P12276>(ctx, a, blocks); break;
            case 16368: rc = launch_fwd_small<512, kLdsCap, P16368>(ctx, a, blocks); break;
            default:
                if (a.N <= 4096) rc = launch_fwd_small<256, 4096, PlanRT>(ctx, a, blocks);
                else if (a.N <= 8192) rc = launch_fwd_small<256, 8192, PlanRT>(ctx, a, blocks);
                else rc = launch_fwd_small<512, kLdsCap, PlanRT>(ctx, a, blocks);
        }
        if (rc) return rc;
    } else {
        const bool ct = fourstep_341(a.N, a.fs);
        if (!ct && !make_fourstep(a.N, a.fs)) {
            set_error("unsupported transform length " + std::to_string(a.N));
            return UGSDR_ERR_UNSUPPORTED;
        }
        UGSDR_HIP_TRY(ctx->aux2.reserve((size_t)blocks * a.N * sizeof(float2)));
        a.scratch = ctx->aux2.as<float2>();
        if (!ct) rc = launch_fwd_large<PlanRT, PlanRT>(ctx, a, blocks);
        else switch (a.fs.n1) {
            case 60: rc = launch_fwd_large<P60, P341>(ctx, a, blocks); break;
            case 112: rc = launch_fwd_large<P112, P341>(ctx, a, blocks); break;
            case 120: rc = launch_fwd_large<P120, P341>(ctx, a, blocks); break;
            case 180: rc = launch_fwd_large<P180, P341>(ctx, a, blocks); break;
            default: rc = launch_fwd_large<P240, P341>(ctx, a, blocks); break;
        }
        if (rc) return rc;
    }
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

template <int T, int CAP, class P>
static int launch_corr_small(ugsdr_hip_ctx* ctx, const CorrArgs& c, int grid) {
    const size_t lds = small_lds_bytes(c.N, T);
    UGSDR_HIP_TRY(set_lds(corr_small<T, CAP, P>, lds));
    hipLaunchKernelGGL((corr_small<T, CAP, P>), dim3(grid), dim3(T), lds, ctx->stream, c);
    return UGSDR_OK;
}

template <class P1, class P2>
static int launch_corr_large(ugsdr_hip_ctx* ctx, const CorrArgs& c, int grid) {
    constexpr int T = 512;
    const size_t lds = large_lds_bytes(c.fs);
    UGSDR_HIP_TRY(set_lds(corr_large<T, P1, P2>, lds));
    hipLaunchKernelGGL((corr_large<T, P1, P2>), dim3(grid), dim3(T), lds, ctx->stream, c);
    return UGSDR_OK;
}

int replica_spectrum(ugsdr_hip_ctx* ctx, const int8_t* chips, int64_t n_chips, int64_t repeats, const float* samples,
                     int64_t n_fft, float2* out) {
    FwdArgs a{};
    a.N = (int)n_fft;
    a.F = 1;
    a.out = out;
    if (chips) {
        UGSDR_HIP_TRY(ctx->aux.reserve((size_t)n_chips));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->aux.ptr, chips, (size_t)n_chips, hipMemcpyHostToDevice, ctx->stream));
        a.mode = 1;
        a.chips = ctx->aux.as<signed char>();
        a.n_chips = n_chips;
        a.len = n_chips * repeats;
    } else {
        UGSDR_HIP_TRY(ctx->aux.reserve((size_t)n_fft * sizeof(float)));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->aux.ptr, samples, (size_t)n_fft * sizeof(float), hipMemcpyHostToDevice,
                                     ctx->stream));
        a.mode = 2;
        a.samples = ctx->aux.as<float>();
    }
    int rc = launch_forward(ctx, a, 1);
    if (rc) return rc;
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

// Doppler bins -> (class, integer shift)
static void doppler_classes(const double* dop, int64_t D, int64_t N, double fs, std::vector<double>& cls_frac,
                            std::vector<int>& cls, std::vector<int>& shift) {
    cls_frac.clear();
    cls.resize(D);
    shift.resize(D);
    for (int64_t i = 0; i < D; ++i) {
        const double s = dop[i] * (double)N / fs;
        double m = std::floor(s);
        double phi = s - m;
        if (phi > 1.0 - 1e-9) { phi = 0.0; m += 1.0; }
        if (phi < 1e-9) phi = 0.0;
        int c = -1;
        for (size_t j = 0; j < cls_frac.size(); ++j)
            if (std::fabs(cls_frac[j] - phi) < 1e-9) { c = (int)j; break; }
        if (c < 0) { c = (int)cls_frac.size(); cls_frac.push_back(phi); }
        cls[i] = c;
        long long mm = (long long)m % (long long)N;
        if (mm < 0) mm += N;
        shift[i] = (int)mm;
    }
}

static int64_t gcd64(int64_t a, int64_t b) { while (b) { int64_t t = a % b; a = b; b = t; } return a; }

int acquire_impl(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, ugsdr_acq_result* out, float* surfaces,
                 float* peaks) {
    UGSDR_CHECK_ARG(r && out, "null request/result");
    UGSDR_CHECK_ARG(r->iq && r->n_samples > 0, "empty signal");
    UGSDR_CHECK_ARG(r->fmt >= UGSDR_FMT_CI8 && r->fmt <= UGSDR_FMT_R8, "unknown sample format");
    UGSDR_CHECK_ARG(r->n_fft > 0 && r->n_blocks > 0 && r->surface_len > 0, "bad sizes");
    UGSDR_CHECK_ARG(r->n_doppler > 0 && r->doppler, "no Doppler bins");
    UGSDR_CHECK_ARG(r->n_prn > 0 && r->replica_keys, "no replicas");
    UGSDR_CHECK_ARG(r->fs_in > 0 && r->fs_acq > 0, "bad sampling rate");
    if (r->n_fft % r->surface_len) {
        set_error("Signal size is not divided by block size");
        return UGSDR_ERR_SIZE;
    }
    if (!r->reshape && r->surface_len != r->n_fft) {
        set_error("surface_len must equal n_fft without reshape");
        return UGSDR_ERR_ARG;
    }
    if (r->precision != UGSDR_PREC_FP32) return acquire_reduced(ctx, r, out, surfaces, peaks);
    const int64_t N = r->n_fft, K = r->n_blocks, L = r->surface_len, D = r->n_doppler, P = r->n_prn;
    if (N > (1 << 22) || D * P > (1 << 30)) {
        set_error("search too large");
        return UGSDR_ERR_UNSUPPORTED;
    }
    // ---- replicas
    std::vector<const float2*> specs(P);
    for (int64_t p = 0; p < P; ++p) {
        auto it = ctx->replicas.find(r->replica_keys[p]);
        if (it == ctx->replicas.end()) {
            set_error("unknown replica key");
            return UGSDR_ERR_STATE;
        }
        if (it->second.n_fft != N || it->second.precision != UGSDR_PREC_FP32) {
            set_error("replica length/precision does not match the request");
            return UGSDR_ERR_ARG;
        }
        specs[p] = static_cast<const float2*>(it->second.spec);
    }
    // ---- front end: unpack + mix + resample
    const int64_t fs_in = (int64_t)std::llround(r->fs_in), fs_acq = (int64_t)std::llround(r->fs_acq);
    int64_t up = 1, down = 1;
    if (fs_in != fs_acq) {
        const int64_t g = gcd64(fs_acq, fs_in);
        const int64_t lcm = fs_acq / g * fs_in;
        up = lcm / fs_in;
        down = lcm / fs_acq;
    }
    const int64_t need = N * K;
    const int64_t avail_out = (up == 1 && down == 1) ? r->n_samples : (r->n_samples * up + down - 1) / down;
    if (avail_out < need) {
        set_error("signal shorter than n_blocks * n_fft at fs_acq");
        return UGSDR_ERR_SIZE;
    }
    const void* iq = r->iq;
    const int64_t n_in_used = std::min<int64_t>(r->n_samples, (need * down + up - 1) / up + 1);
    if (!r->iq_on_device) {
        const size_t bps = r->fmt == UGSDR_FMT_CI8 ? 2 : r->fmt == UGSDR_FMT_CI16 ? 4 : r->fmt == UGSDR_FMT_CF32 ? 8 : 1;
        UGSDR_HIP_TRY(ctx->iq.reserve(n_in_used * bps));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->iq.ptr, r->iq, n_in_used * bps, hipMemcpyHostToDevice, ctx->stream));
        iq = ctx->iq.ptr;
    }
    UGSDR_HIP_TRY(ctx->sig.reserve((size_t)need * sizeof(float2)));
    {
        FrontArgs fa{};
        fa.iq = iq;
        fa.fmt = r->fmt;
        fa.n_in = n_in_used;
        fa.n_out = need;
        double f = r->mix_frequency;
        if (f < 0) f = r->fs_in + f;  // Mixer::Translate folding (mixer.hpp:46-47)
        double ph = std::fmod(r->mix_phase, 2 * M_PI);
        if (ph < 0) ph += 2 * M_PI;
        fa.apply_tone = (f != 0.0 || ph != 0.0) ? 1 : 0;
        fa.f_turns = f / r->fs_in;
        fa.ph_turns = ph / (2 * M_PI);
        fa.up = up;
        fa.down = down;
        fa.out = ctx->sig.as<float2>();
        const int blocks = (int)std::min<int64_t>((need + 255) / 256, 4096);
        hipLaunchKernelGGL(front_kernel, dim3(blocks), dim3(256), 0, ctx->stream, fa);
        UGSDR_HIP_TRY(hipGetLastError());
    }
    // ---- Doppler classes and tables
    std::vector<double> cls_frac;
    std::vector<int> cls, shift;
    doppler_classes(r->doppler, D, N, r->fs_acq, cls_frac, cls, shift);
    const int F = (int)cls_frac.size();
    const size_t tab_bytes = cls_frac.size() * 8 + (size_t)D * 8 + (size_t)D * 4 * 2 + (size_t)P * 8 + 64;
    UGSDR_HIP_TRY(ctx->tables.reserve(tab_bytes));
    std::vector<char> host_tab(tab_bytes);
    char* hp = host_tab.data();
    size_t off_frac = 0, off_dop = off_frac + cls_frac.size() * 8, off_cls = off_dop + D * 8,
           off_shift = off_cls + D * 4, off_specs = (off_shift + D * 4 + 7) / 8 * 8;
    memcpy(hp + off_frac, cls_frac.data(), cls_frac.size() * 8);
    memcpy(hp + off_dop, r->doppler, D * 8);
    memcpy(hp + off_cls, cls.data(), D * 4);
    memcpy(hp + off_shift, shift.data(), D * 4);
    memcpy(hp + off_specs, specs.data(), P * 8);
    UGSDR_HIP_TRY(hipMemcpyAsync(ctx->tables.ptr, hp, off_specs + P * 8, hipMemcpyHostToDevice, ctx->stream));
    char* dp = ctx->tables.as<char>();
    // ---- forward transforms of each block for each class
    UGSDR_HIP_TRY(ctx->classes.reserve((size_t)K * F * N * sizeof(float2)));
    {
        FwdArgs a{};
        a.N = (int)N;
        a.F = F;
        a.sig = ctx->sig.as<float2>();
        a.class_frac = reinterpret_cast<const double*>(dp + off_frac);
        a.out = ctx->classes.as<float2>();
        a.mode = 0;
        int rc = launch_forward(ctx, a, (int)(K * F));
        if (rc) return rc;
    }
    // ---- correlation rows
    const int rows = (int)(P * D);
    UGSDR_HIP_TRY(ctx->rows.reserve((size_t)rows * sizeof(RowStat)));
    const bool keep_surf = surfaces || peaks;
    if (keep_surf) UGSDR_HIP_TRY(ctx->surf.reserve((size_t)rows * L * sizeof(float)));
    CorrArgs c{};
    c.N = (int)N;
    c.K = (int)K;
    c.F = F;
    c.L = (int)L;
    c.D = (int)D;
    c.P = (int)P;
    c.fold_mode = (r->reshape && !r->coherent) ? 1 : 0;
    c.rows = rows;
    c.X = ctx->classes.as<float2>();
    c.specs = reinterpret_cast<const float2* const*>(dp + off_specs);
    c.cls = reinterpret_cast<const int*>(dp + off_cls);
    c.shift = reinterpret_cast<const int*>(dp + off_shift);
    c.rowstats = ctx->rows.as<RowStat>();
    c.surfaces = keep_surf ? ctx->surf.as<float>() : nullptr;
    c.inv_n = (float)(1.0 / (double)N);
    const int cus = cu_count(ctx->device);
    if (N <= kLdsCap) {
        if (!make_plan((int)N, c.plan)) {
            set_error("unsupported transform length " + std::to_string(N));
            return UGSDR_ERR_UNSUPPORTED;
        }
        const int grid = rows;
        int rc;
        switch (N) {
            case 4092: rc = launch_corr_small<256, 4096, P4092>(ctx, c, grid); break;
            case 8184: rc = launch_corr_small<256, 8192, P8184>(ctx, c, grid); break;
            case 12276: rc = launch_corr_small<512, kLdsCap, P12276>(ctx, c, grid); break;
            case 16368: rc = launch_corr_small<512, kLdsCap, P16368>(ctx, c, grid); break;
            default:
                if (N <= 4096) rc = launch_corr_small<256, 4096, PlanRT>(ctx, c, grid);
                else if (N <= 8192) rc = launch_corr_small<256, 8192, PlanRT>(ctx, c, grid);
                else rc = launch_corr_small<512, kLdsCap, PlanRT>(ctx, c, grid);
        }
        if (rc) return rc;
    } else {
        const bool ct = fourstep_341((int)N, c.fs);
        if (!ct && !make_fourstep((int)N, c.fs)) {
            set_error("unsupported transform length " + std::to_string(N));
            return UGSDR_ERR_UNSUPPORTED;
        }
        const int grid = std::min(rows, cus);
        const size_t per_wg = (2 * (size_t)N + (L + 1) / 2) * sizeof(float2);
        UGSDR_HIP_TRY(ctx->scratch.reserve(per_wg * grid));
        c.scratch = ctx->scratch.as<float2>();
        int rc;
        if (!ct) rc = launch_corr_large<PlanRT, PlanRT>(ctx, c, grid);
        else switch (c.fs.n1) {
            case 60: rc = launch_corr_large<P60, P341>(ctx, c, grid); break;
            case 112: rc = launch_corr_large<P112, P341>(ctx, c, grid); break;
            case 120: rc = launch_corr_large<P120, P341>(ctx, c, grid); break;
            case 180: rc = launch_corr_large<P180, P341>(ctx, c, grid); break;
            default: rc = launch_corr_large<P240, P341>(ctx, c, grid); break;
        }
        if (rc) return rc;
    }
    UGSDR_HIP_TRY(hipGetLastError());
    // ---- finalize
    UGSDR_HIP_TRY(ctx->results.reserve((size_t)P * sizeof(ugsdr_acq_result) + (peaks ? (size_t)P * L * 4 : 0)));
    FinalArgs fa{};
    fa.rowstats = ctx->rows.as<RowStat>();
    fa.doppler = reinterpret_cast<const double*>(dp + off_dop);
    fa.D = (int)D;
    fa.P = (int)P;
    fa.ratio = r->fs_in / r->fs_acq;
    fa.intermediate_frequency = r->intermediate_frequency;
    fa.threshold = r->threshold;
    fa.reshape = r->reshape;
    fa.out = ctx->results.as<ugsdr_acq_result>();
    fa.surfaces = keep_surf ? ctx->surf.as<float>() : nullptr;
    fa.peaks = peaks ? reinterpret_cast<float*>(ctx->results.as<char>() + (size_t)P * sizeof(ugsdr_acq_result))
                     : nullptr;
    fa.L = (int)L;
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)P), dim3(256), 0, ctx->stream, fa);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, fa.out, (size_t)P * sizeof(ugsdr_acq_result), hipMemcpyDeviceToHost,
                                 ctx->stream));
    if (peaks)
        UGSDR_HIP_TRY(hipMemcpyAsync(peaks, fa.peaks, (size_t)P * L * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (surfaces)
        UGSDR_HIP_TRY(hipMemcpyAsync(surfaces, ctx->surf.ptr, (size_t)rows * L * 4, hipMemcpyDeviceToHost,
                                     ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

}  // namespace ugsdr_hip

extern "C" int ugsdr_hip_acquire(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* req, ugsdr_acq_result* out,
                                 float* surfaces, float* peaks) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return ugsdr_hip::acquire_impl(ctx, req, out, surfaces, peaks);
}
