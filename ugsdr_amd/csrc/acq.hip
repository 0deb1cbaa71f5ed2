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
#include <string>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "plan.hpp"
#include "internal.hpp"
#include "acq.hpp"
#include "formats.hpp"
#include "corr_common.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

// ------------------------------------------------------------------ front end
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
            float2 v = load_sample(a.src, o);
            if (a.apply_tone) v = cmul(v, phasor_turns(fma((double)o, a.f_turns, a.ph_turns)));
            a.out[o] = v;
        } else {
            double sr = 0.0, si = 0.0;
            const int64_t u0 = o * a.down;
            for (int64_t u = u0; u < u0 + a.down; ++u) {
                const int64_t i = u / a.up;
                if (i >= a.n_in) break;
                float2 v = load_sample(a.src, i);
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
    const short2* sig16;         // mode 0: int16 I/Q read in place of sig when set
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
        const int64_t i = (int64_t)k * a.N + n;
        float2 x;
        if (a.sig16) {
            const short2 v = a.sig16[i];
            x = make_float2((float)v.x, (float)v.y);
        } else {
            x = a.sig[i];
        }
        const double phi = a.class_frac[c];
        if (phi == 0.0) return x;
        // the reciprocal is loop-invariant in the callers' sample loops
        return cmul(x, phasor_turns(-phi * ((double)n * (1.0 / (double)a.N))));
    }
    if (a.mode == 1) {
        const int64_t m = ((n * a.len) / a.N) % a.len;
        return make_float2((float)a.chips[m % a.n_chips], 0.f);
    }
    return make_float2(a.samples[n], 0.f);
}

// waves per SIMD to budget registers for: 3 for N <= 4096 (three 256-thread
// workgroups per CU), else 2 (LDS holds one or two transforms per CU)
constexpr int waves_for(int CAP) { return CAP <= 4096 ? 3 : 2; }
// a workgroup of T threads puts T/256 waves on each SIMD
constexpr int waves_for_t(int T, int CAP) { return T / 256 > waves_for(CAP) ? T / 256 : waves_for(CAP); }

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
// Sequences sit in LDS at an odd stride (conflict-free strided access of
// 8-byte elements); batches are balanced over their count; every global
// loop keeps several loads in flight per thread.  sink_pre(k) is read for
// output k before sink(k, value, pre) is called (so a read-modify-write of
// the output batches its reads).
template <int T, bool INV, class P1, class P2, class Src, class SinkPre, class Sink>
__device__ void fourstep(float2* lds, const FourStep& f, const float2* lo1, const float2* hi1, const float2* lo2,
                         const float2* hi2, const float2* loN, const float2* hiN, float2* G, Src src,
                         SinkPre sink_pre, Sink sink, int tid0) {
    const int N1 = P1::kCompileTime ? P1::N : f.n1;
    const int N2 = P2::kCompileTime ? P2::N : f.n2;
    const int ld1 = N1 | 1, ld2 = N2 | 1;
    const int nbat1 = (N2 + max(1, kLdsCap / ld1) - 1) / max(1, kLdsCap / ld1);
    const int cb = (N2 + nbat1 - 1) / nbat1;
    for (int c0 = 0; c0 < N2; c0 += cb) {
        const int tid = opaque(tid0);
        const int nb = min(cb, N2 - c0);
        walk_cols<8>(
            tid, T, nb, N1, [&](int t, int n1) { return src((int64_t)N2 * n1 + c0 + t); },
            [&](int t, int n1, float2 v) { lds[t * ld1 + n1] = v; });
        __syncthreads();
        Fft<T, kLdsCap, INV, P1>::run(lds, f.p1, nb, lo1, hi1, tid, ld1);
        walk_cols<4>(
            tid, T, nb, N1, [&](int t, int k1) { return lds[t * ld1 + k1]; },
            [&](int t, int k1, float2 v) {
                const int n2 = c0 + t;
                G[(int64_t)k1 * N2 + n2] = cmul(v, tw_get<INV>(loN, hiN, n2 * k1));
            });
        __syncthreads();
    }
    const int nbat2 = (N1 + max(1, kLdsCap / ld2) - 1) / max(1, kLdsCap / ld2);
    const int rb = (N1 + nbat2 - 1) / nbat2;
    for (int r0 = 0; r0 < N1; r0 += rb) {
        const int tid = opaque(tid0);
        const int nb = min(rb, N1 - r0);
        walk_cols<8>(
            tid, T, N2, nb, [&](int n2, int t) { return G[(int64_t)(r0 + t) * N2 + n2]; },
            [&](int n2, int t, float2 v) { lds[t * ld2 + n2] = v; });
        __syncthreads();
        Fft<T, kLdsCap, INV, P2>::run(lds, f.p2, nb, lo2, hi2, tid, ld2);
        walk_cols<8>(
            tid, T, nb, N2,
            [&](int t, int k2) {
                const int64_t k = (int64_t)(r0 + t) + (int64_t)N1 * k2;
                return make_float3(lds[t * ld2 + k2].x, lds[t * ld2 + k2].y, sink_pre(k));
            },
            [&](int t, int k2, float3 v) {
                sink((int64_t)(r0 + t) + (int64_t)N1 * k2, make_float2(v.x, v.y), v.z);
            });
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
        [&](int64_t n) { return fwd_source(a, blk, n); }, [](int64_t) { return 0.f; },
        [&](int64_t k, float2 v, float) {
            if (conj_out) v.y = -v.y;
            o[k] = v;
        },
        tid);
}

// The four-step forward transform split over workgroups (both factors with
// compile-time plans): PH 0 runs one batch of `batch` columns of one
// transform (N1-point DFTs and the twiddle w_N^(n2 k1), into rows k1 of the
// transform's scratch G), PH 1 one batch of `batch` rows (N2-point DFTs,
// outputs in natural order).  The two launches carry blocks x batches small
// workgroups each, so a handful of transforms (a search's K x F class
// spectra) spreads over the device instead of running one transform per
// workgroup.
constexpr int kSplitCap = 4096;   // complex values of a workgroup's batch

template <int T, int PH, class P1, class P2>
__global__ __launch_bounds__(T) void fwd_split(FwdArgs a, int nbatch, int batch) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int N1 = P1::N, N2 = P2::N;
    float2* buf = reinterpret_cast<float2*>(smem);
    float2* lo = buf + kSplitCap;
    float2* hi = lo + kTwLo;
    const int tid = threadIdx.x;
    const int blk = blockIdx.x / nbatch, bat = blockIdx.x - blk * nbatch;
    float2* G = a.scratch + (int64_t)blk * a.N;
    if constexpr (PH == 0) {
        constexpr int ld1 = N1 | 1;
        float2* loN = lo + tw_len(N1);
        float2* hiN = loN + kTwLo;
        tw_build(lo, hi, N1, tid, T);
        tw_build(loN, hiN, a.N, tid, T);
        const int c0 = bat * batch, nb = min(batch, N2 - c0);
        walk_cols<8>(
            tid, T, nb, N1, [&](int t, int n1) { return fwd_source(a, blk, (int64_t)N2 * n1 + c0 + t); },
            [&](int t, int n1, float2 v) { buf[t * ld1 + n1] = v; });
        __syncthreads();
        Fft<T, kSplitCap, false, P1>::run(buf, a.fs.p1, nb, lo, hi, tid, ld1);
        walk_cols<4>(
            tid, T, nb, N1, [&](int t, int k1) { return buf[t * ld1 + k1]; },
            [&](int t, int k1, float2 v) {
                const int n2 = c0 + t;
                G[(int64_t)k1 * N2 + n2] = cmul(v, tw_get<false>(loN, hiN, n2 * k1));
            });
    } else {
        constexpr int ld2 = N2 | 1;
        tw_build(lo, hi, N2, tid, T);
        const int r0 = bat * batch, nb = min(batch, N1 - r0);
        walk_cols<8>(
            tid, T, N2, nb, [&](int n2, int t) { return G[(int64_t)(r0 + t) * N2 + n2]; },
            [&](int n2, int t, float2 v) { buf[t * ld2 + n2] = v; });
        __syncthreads();
        Fft<T, kSplitCap, false, P2>::run(buf, a.fs.p2, nb, lo, hi, tid, ld2);
        float2* o = a.out + (int64_t)blk * a.N;
        const bool conj_out = a.mode != 0;
        walk_cols<8>(
            tid, T, nb, N2, [&](int t, int k2) { return buf[t * ld2 + k2]; },
            [&](int t, int k2, float2 v) {
                if (conj_out) v.y = -v.y;
                o[(int64_t)(r0 + t) + (int64_t)N1 * k2] = v;
            });
    }
}

template <int T, int CAP, class P>
__global__ __launch_bounds__(T, waves_for_t(T, CAP)) void corr_small(CorrArgs a) {
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
        const ReqDesc q = a.reqs[find_req(a, row)];
        const int local = row - q.row_start;
        const int p = local / q.D, di = local - p * q.D;
        const int cls = q.cls[di], sh = q.shift[di];
        const gptr<float2> C = as_global(q.specs[p]);
        float S[SR];
#pragma unroll
        for (int i = 0; i < SR; ++i) S[i] = 0.f;
        for (int k = 0; k < q.K; ++k) {
            const int tq = opaque(tid);
            const gptr<float2> Xk = as_global(q.X + ((int64_t)k * q.F + cls) * N);
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
                            S[i] += cabs_hw(v);
                        }
                    }
                } else {
                    for (int j = tq; j < N; j += T) {
                        const float2 v = buf[j];
                        buf[j].x = cabs_hw(v);
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
                        S[i] += cabs_hw(z);
                    }
                }
            }
            __syncthreads();
        }
        row_stats_reg<T, SR>(S, L, red, tid, q.rowstats + local, a.inv_n);
        if (q.surfaces) {
            float* o = q.surfaces + (int64_t)local * L;
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
__global__ __launch_bounds__(T, (T / 256 > 2 ? T / 256 : 2)) void corr_large(CorrArgs a) {
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
        const ReqDesc q = a.reqs[find_req(a, row)];
        const int local = row - q.row_start;
        const int p = local / q.D, di = local - p * q.D;
        const int cls = q.cls[di], sh = q.shift[di];
        const gptr<float2> C = as_global(q.specs[p]);
        // the last block's surface values go straight into the thread's share
        // of the row statistics (top two, FP64 sum and sum of squares) and,
        // when asked for, the caller's surface: the row's running surface S
        // is neither written for it nor read back
        float v1 = -1.0f, v2 = -1.0f;
        int i1 = 0x7fffffff;
        double sum = 0.0, sumsq = 0.0;
        float* surf = q.surfaces ? q.surfaces + (int64_t)local * L : nullptr;
        const float inv_n = a.inv_n;
        auto finish = [&](int64_t j, float val) {
            top2_merge(v1, i1, v2, val, (int)j, -1.0f);
            sum += (double)val;
            sumsq += (double)val * (double)val;
            if (surf) surf[j] = val * inv_n;
        };
        for (int k = 0; k < q.K; ++k) {
            const int tq = opaque(tid);
            const gptr<float2> Xk = as_global(q.X + ((int64_t)k * q.F + cls) * N);
            const int fm = a.fold_mode;
            const bool first = k == 0, last = k == q.K - 1;
            fourstep<T, true, P1, P2>(
                Ld.buf, a.fs, Ld.lo1, Ld.hi1, Ld.lo2, Ld.hi2, Ld.loN, Ld.hiN, G,
                [&](int64_t n) {
                    int64_t src = n + sh;
                    src -= src >= N ? N : 0;
                    return cmul(Xk[src], C[n]);
                },
                [&](int64_t kk) { return R == 1 && !first ? S[kk] : 0.f; },
                [&](int64_t kk, float2 v, float pre) {
                    // unfolded surface: |.| straight into the block sum
                    // (each output index is written by one thread)
                    if (R == 1) {
                        const float val = pre + cabs_hw(v);
                        if (last) finish(kk, val);
                        else S[kk] = val;
                    } else if (fm == 0) {
                        A[kk] = cabs_hw(v);
                    } else {
                        B[kk] = v;
                    }
                },
                tq);
            if (R == 1) {
                __syncthreads();
                continue;
            }
            for (int j = tq; j < L; j += T) {
                float acc;
                if (fm == 0) {
                    acc = A[j];
                    for (int r = 1; r < R; ++r) acc += A[j + (int64_t)r * L];
                } else {
                    float2 z = B[j];
                    for (int r = 1; r < R; ++r) z = cadd(z, B[j + (int64_t)r * L]);
                    acc = cabs_hw(z);
                }
                const float val = first ? acc : S[j] + acc;
                if (last) finish(j, val);
                else S[j] = val;
            }
            __syncthreads();
        }
        // sum of squared deviations from the FP64 sums (a positive surface
        // whose spread is comparable to its mean: no precision to speak of lost)
        block_top2<T>(v1, i1, v2, Ld.red, tid);
        sum = block_sum<T>(sum, Ld.red + 64, tid);
        sumsq = block_sum<T>(sumsq, Ld.red + 64, tid);
        const double mean = sum / (double)L;
        const double dev = sumsq - sum * mean;
        stats_write<T>(v1, i1, v2, mean, dev > 0.0 ? dev : 0.0, L, tid, q.rowstats + local, a.inv_n);
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

// The selection runs in bin order on thread 0 (strict <, runner-up as it
// goes); the row statistics reach it through LDS, 256 bins at a time,
// loaded by the whole workgroup.
__global__ __launch_bounds__(256) void finalize_kernel(FinalArgs a) {
    const int p = blockIdx.x;
    __shared__ int sel;
    __shared__ RowStat rs[256];
    ugsdr_acq_result r;
    double runner = 0.0;
    if (threadIdx.x == 0) {
        r.level = 0.0;
        r.sigma = 1.0;
        r.mean = 0.0;
        r.stddev = 0.0;
        r.doppler = 0.0;
        r.code_offset = 0.0;
        r.runner_up_code = 0.0;
        r.index = 0;
        r.doppler_bin = -1;
    }
    for (int d0 = 0; d0 < a.D; d0 += 256) {
        const int nd = min(256, a.D - d0);
        __syncthreads();
        if ((int)threadIdx.x < nd) rs[threadIdx.x] = a.rowstats[(int64_t)p * a.D + d0 + threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 0; i < nd; ++i) {
                const int di = d0 + i;
                const RowStat s = rs[i];
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
        }
    }
    if (threadIdx.x == 0) {
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
    return set_lds_limit(reinterpret_cast<const void*>(kernel), bytes);
}

int device_cu_count(int device) {
    static int cached[64] = {0};
    if (device < 0 || device >= 64) return 256;
    if (!cached[device]) {
        hipDeviceProp_t prop;
        int n = 0;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) n = prop.multiProcessorCount;
        cached[device] = n > 0 ? n : 256;
    }
    return cached[device];
}

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
using P128 = PlanCT<16, 8>;
using P256 = PlanCT<16, 16>;
using P4200 = PlanCT<8, 7, 5, 5, 3>;  // 1 ms at 4.2 MHz (the BbpDdc recordings' rate rule)
// first factor of the forward four-step split 16368 = 48 x 341
using P48 = PlanCT<16, 3>;
using P152 = PlanCT<19, 8>;
using P175 = PlanCT<7, 5, 5>;

template <int T, int CAP, class P>
static int launch_fwd_small(FwdArgs& a, int blocks, hipStream_t st) {
    const size_t lds = small_lds_bytes(a.N, T);
    UGSDR_HIP_TRY(set_lds(fwd_small<T, CAP, P>, lds));
    hipLaunchKernelGGL((fwd_small<T, CAP, P>), dim3(blocks), dim3(T), lds, st, a);
    return UGSDR_OK;
}

template <class P1, class P2>
static int launch_fwd_large(FwdArgs& a, int blocks, hipStream_t st) {
    constexpr int T = 512;
    const size_t lds = large_lds_bytes(a.fs);
    UGSDR_HIP_TRY(set_lds(fwd_large<T, P1, P2>, lds));
    hipLaunchKernelGGL((fwd_large<T, P1, P2>), dim3(blocks), dim3(T), lds, st, a);
    return UGSDR_OK;
}

// batches of at most 16 columns and N1 / 4 rows, so that even one or two
// transforms spread over tens of workgroups
template <class P1, class P2>
static int launch_fwd_split(FwdArgs& a, int blocks, hipStream_t st, DevBuf& scratch) {
    constexpr int T = 256, N1 = P1::N, N2 = P2::N;
    UGSDR_HIP_TRY(scratch.reserve((size_t)blocks * a.N * sizeof(float2)));
    a.scratch = scratch.as<float2>();
    static_assert((N1 | 1) <= kSplitCap && (N2 | 1) <= kSplitCap, "a sequence exceeds a split batch");
    const int cb = std::min({kSplitCap / (N1 | 1), N2, 16});
    const int rb = std::min(kSplitCap / (N2 | 1), std::max(1, N1 / 4));
    const int nb1 = (N2 + cb - 1) / cb, nb2 = (N1 + rb - 1) / rb;
    const size_t lds0 = (size_t)(kSplitCap + tw_len(N1) + tw_len(a.N)) * sizeof(float2);
    const size_t lds1 = (size_t)(kSplitCap + tw_len(N2)) * sizeof(float2);
    UGSDR_HIP_TRY(set_lds(fwd_split<T, 0, P1, P2>, lds0));
    UGSDR_HIP_TRY(set_lds(fwd_split<T, 1, P1, P2>, lds1));
    hipLaunchKernelGGL((fwd_split<T, 0, P1, P2>), dim3(blocks * nb1), dim3(T), lds0, st, a, nb1, cb);
    hipLaunchKernelGGL((fwd_split<T, 1, P1, P2>), dim3(blocks * nb2), dim3(T), lds1, st, a, nb2, rb);
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

// four-step splits whose two factors both have compile-time plans: N1 * 341
// as above, the powers of two 2^15 = 128 * 256 and 2^16 = 256 * 256, and
// 26600 = 152 * 175 (4 ms at 6.65 MHz, the BbpDdc recordings' acquisition
// rate; the same split the runtime planner picks)
static bool fourstep_ct(int N, FourStep& f) {
    if (fourstep_341(N, f)) return true;
    if (N == 26600) {
        f.n = N;
        f.n1 = 152;
        f.n2 = 175;
        make_plan(f.n1, f.p1);
        make_plan(f.n2, f.p2);
        return true;
    }
    if (N != 32768 && N != 65536) return false;
    f.n = N;
    f.n1 = N / 256;
    f.n2 = 256;
    make_plan(f.n1, f.p1);
    make_plan(f.n2, f.p2);
    return true;
}

static bool plan_transform(int N, bool& small, RadixPlan& plan, FourStep& fs) {
    small = N <= kLdsCap;
    if (small) return make_plan(N, plan);
    return fourstep_ct(N, fs) || make_fourstep(N, fs);
}

// forward transform(s): `blocks` transforms of length a.N written to a.out
static int launch_forward(FwdArgs& a, int blocks, hipStream_t st, DevBuf& scratch) {
    int rc = UGSDR_OK;
    bool small;
    if (!plan_transform(a.N, small, a.plan, a.fs)) {
        set_error("unsupported transform length " + std::to_string(a.N));
        return UGSDR_ERR_UNSUPPORTED;
    }
    if (small) {
        // 16368 = 48 x 341 (the wide L1 C/A search, one workgroup's LDS per
        // transform): columns and rows over many small workgroups, as the
        // four-step lengths (measured 18.3 against 24.2 us for 40 transforms;
        // at 4092 with two transforms the two launches cost more than they save)
        if (a.N == 16368) {
            a.fs.n = a.N;
            a.fs.n1 = 48;
            a.fs.n2 = 341;
        }
        switch (a.N) {
            case 4092: rc = launch_fwd_small<512, 4096, P4092>(a, blocks, st); break;
            case 8184: rc = launch_fwd_small<512, 8192, P8184>(a, blocks, st); break;
            case 4200: rc = launch_fwd_small<512, 8192, P4200>(a, blocks, st); break;
            case 12276: rc = launch_fwd_small<1024, kLdsCap, P12276>(a, blocks, st); break;
            case 16368: rc = launch_fwd_split<P48, P341>(a, blocks, st, scratch); break;
            default:
                if (a.N <= 4096) rc = launch_fwd_small<256, 4096, PlanRT>(a, blocks, st);
                else if (a.N <= 8192) rc = launch_fwd_small<256, 8192, PlanRT>(a, blocks, st);
                else rc = launch_fwd_small<512, kLdsCap, PlanRT>(a, blocks, st);
        }
    } else {
        UGSDR_HIP_TRY(scratch.reserve((size_t)blocks * a.N * sizeof(float2)));
        a.scratch = scratch.as<float2>();
        // compile-time splits: columns and rows over many small workgroups;
        // runtime splits: one workgroup per transform
        FourStep probe;
        const bool ct = fourstep_ct(a.N, probe);
        if (!ct) rc = launch_fwd_large<PlanRT, PlanRT>(a, blocks, st);
        else if (a.fs.n2 == 175) rc = launch_fwd_split<P152, P175>(a, blocks, st, scratch);
        else if (a.fs.n2 == 256) rc = a.fs.n1 == 128 ? launch_fwd_split<P128, P256>(a, blocks, st, scratch)
                                                       : launch_fwd_split<P256, P256>(a, blocks, st, scratch);
        else switch (a.fs.n1) {
            case 60: rc = launch_fwd_split<P60, P341>(a, blocks, st, scratch); break;
            case 112: rc = launch_fwd_split<P112, P341>(a, blocks, st, scratch); break;
            case 120: rc = launch_fwd_split<P120, P341>(a, blocks, st, scratch); break;
            case 180: rc = launch_fwd_split<P180, P341>(a, blocks, st, scratch); break;
            default: rc = launch_fwd_split<P240, P341>(a, blocks, st, scratch); break;
        }
    }
    if (rc) return rc;
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

template <int T, int CAP, class P>
static int launch_corr_small(const CorrArgs& c, int grid, hipStream_t st) {
    const size_t lds = small_lds_bytes(c.N, T);
    UGSDR_HIP_TRY(set_lds(corr_small<T, CAP, P>, lds));
    hipLaunchKernelGGL((corr_small<T, CAP, P>), dim3(grid), dim3(T), lds, st, c);
    return UGSDR_OK;
}

template <int T, class P1, class P2>
static int launch_corr_large(const CorrArgs& c, int grid, hipStream_t st) {
    const size_t lds = large_lds_bytes(c.fs);
    UGSDR_HIP_TRY(set_lds(corr_large<T, P1, P2>, lds));
    hipLaunchKernelGGL((corr_large<T, P1, P2>), dim3(grid), dim3(T), lds, st, c);
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
    int rc = launch_forward(a, 1, ctx->stream, ctx->aux2);
    if (rc) return rc;
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

// Doppler bins -> (class, integer shift): d N / fs = m + phi
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

int64_t AcqPlan::footprint(bool surfaces) const {
    int64_t b = in_bytes + need * 8 + K * F * N * 8 + (int64_t)table_bytes + P * D * 32 +
                P * (int64_t)sizeof(ugsdr_acq_result) + P * L * 4;
    if (Tsub > 1) b += K * Tsub * F * N * 8;
    if (fwd_scratch_needed()) b += K * Tsub * F * N * 8;
    if (ucode) b += D * K * N * (precision == UGSDR_PREC_FP32 ? 8 : 4);
    if (surfaces) b += P * D * L * 4;
    return b;
}

// conj of the U-sample boxcar's N-point spectrum, H[k] = sum_{j<U} e^(-j 2 pi jk/N)
// = e^(-j pi (U-1) k/N) sin(pi U k/N) / sin(pi k/N), built in FP64 once per (N, U),
// and the bound max_{k' < N/U} sum_u |H[k' + u N/U]| on |Z| / max |X|
static int boxcar_table(ugsdr_hip_ctx* ctx, int64_t N, int U, const float2** out, float* bound) {
    DevBuf& b = ctx->boxcar[N * 64 + U];
    if (!b.ptr) {
        std::vector<float2> h((size_t)N);
        std::vector<double> habs((size_t)N);
        for (int64_t k = 0; k < N; ++k) {
            double re = (double)U, im = 0.0;
            if (k) {
                const double mag = std::sin(M_PI * (double)U * (double)k / (double)N) / std::sin(M_PI * (double)k / (double)N);
                const double ph = -M_PI * (double)(U - 1) * (double)k / (double)N;
                re = mag * std::cos(ph);
                im = mag * std::sin(ph);
            }
            h[(size_t)k] = make_float2((float)re, (float)-im);
            habs[(size_t)k] = std::hypot(re, im);
        }
        const int64_t Lc = N / U;
        double mx = 0.0;
        for (int64_t k = 0; k < Lc; ++k) {
            double s = 0.0;
            for (int u = 0; u < U; ++u) s += habs[(size_t)(k + Lc * u)];
            mx = std::max(mx, s);
        }
        // the float bound rounds up: a scale derived from it never overflows
        ctx->boxcar_bound[N * 64 + U] = (float)(mx * (1.0 + 1e-6));
        UGSDR_HIP_TRY(b.reserve((size_t)N * sizeof(float2)));
        UGSDR_HIP_TRY(hipMemcpyAsync(b.ptr, h.data(), (size_t)N * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
        UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    *out = b.as<float2>();
    if (bound) *bound = ctx->boxcar_bound[N * 64 + U];
    return UGSDR_OK;
}

int make_acq_plan(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, AcqPlan& pl, bool allow_ucode) {
    pl.precision = r ? r->precision : UGSDR_PREC_FP32;
    UGSDR_CHECK_ARG(r, "null request");
    UGSDR_CHECK_ARG(r->iq && r->n_samples > 0, "empty signal");
    if (int rc = format_geometry(r->fmt, r->fs_in, pl.spm_in, pl.epoch_bytes)) return rc;
    UGSDR_CHECK_ARG(r->n_fft > 0 && r->n_blocks > 0 && r->surface_len > 0, "bad sizes");
    UGSDR_CHECK_ARG(r->n_doppler > 0 && r->doppler, "no Doppler bins");
    UGSDR_CHECK_ARG(r->n_prn > 0 && r->replica_keys, "no replicas");
    UGSDR_CHECK_ARG(r->hypotheses >= 0 && (r->hypotheses <= 1 || r->n_prn % r->hypotheses == 0),
                    "n_prn must be a multiple of the hypothesis count");
    UGSDR_CHECK_ARG(r->fs_in > 0 && r->fs_acq > 0, "bad sampling rate");
    if (r->n_fft % r->surface_len) {
        set_error("Signal size is not divided by block size");
        return UGSDR_ERR_SIZE;
    }
    if (!r->reshape && r->surface_len != r->n_fft) {
        set_error("surface_len must equal n_fft without reshape");
        return UGSDR_ERR_ARG;
    }
    pl.N = r->n_fft;
    pl.K = r->n_blocks;
    pl.L = r->surface_len;
    pl.D = r->n_doppler;
    pl.P = r->n_prn;
    pl.Tsub = 1;
    if (pl.N > (1 << 22) || pl.D * pl.P > (1 << 30) || pl.K > (1 << 20)) {
        set_error("search too large");
        return UGSDR_ERR_UNSUPPORTED;
    }
    pl.reshape = r->reshape;
    pl.fold_mode = (r->reshape && !r->coherent) ? 1 : 0;
    pl.specs.resize(pl.P);
    for (int64_t p = 0; p < pl.P; ++p) {
        auto it = ctx->replicas.find(r->replica_keys[p]);
        if (it == ctx->replicas.end()) {
            set_error("unknown replica key");
            return UGSDR_ERR_STATE;
        }
        if (it->second.n_fft != pl.N || it->second.precision != r->precision) {
            set_error("replica length/precision does not match the request");
            return UGSDR_ERR_ARG;
        }
        pl.specs[p] = static_cast<const float2*>(it->second.spec);
        pl.spec_scale = it->second.scale;
    }
    // Periodic fold (FP32, folded surface, every replica made of whole code
    // periods of surface_len samples): the n_fft-point circular correlation of
    // a replica of T = n_fft / L periods is L-periodic, and its n_fft-point
    // inverse DFT only sees the replica spectrum's bins k = T k', where it is
    // T times the one-period spectrum; so the search folded to L equals an
    // L-point search of the T periods' spectra combined coherently (scaled by
    // T, the folded copies).  The transform shrinks by T.
    {
        bool periodic = r->precision == UGSDR_PREC_FP32 && r->reshape && pl.L < pl.N && pl.N % pl.L == 0 &&
                        pl.N / pl.L <= kMaxPeriods && pl.K * pl.D * (pl.N / pl.L) * 64 < ((int64_t)1 << 31);
        for (int64_t p = 0; p < pl.P && periodic; ++p) {
            const Replica& rep = ctx->replicas.find(r->replica_keys[p])->second;
            periodic = rep.spec_short && rep.period == pl.L;
        }
        if (periodic) {
            pl.Tsub = pl.N / pl.L;
            pl.N = pl.L;
            for (int64_t p = 0; p < pl.P; ++p)
                pl.specs[p] = static_cast<const float2*>(ctx->replicas.find(r->replica_keys[p])->second.spec_short);
            pl.fold_mode = 0;   // the surface is no longer folded: |.| as is
        }
    }
    if (!plan_transform((int)pl.N, pl.small, pl.plan, pl.fs)) {
        set_error("unsupported transform length " + std::to_string(pl.N));
        return UGSDR_ERR_UNSUPPORTED;
    }
    // Upsampled-code search (FP32, unfolded surface, every replica one code
    // period of Lc chips held for U = N / Lc samples each): the replica's
    // spectrum is C[k mod Lc] H[k] with C the chips' Lc-point spectrum and H
    // the U-sample boxcar's, so with k = k' + Lc u and tau = U q + p the
    // row's inverse transform is, per phase p, the Lc-point inverse transform
    // of conj(C) Z_p, where Z_p[k'] = e^(j 2 pi k'p / N) sum_u
    // X_rot[k' + Lc u] conj(H[k' + Lc u]) e^(j 2 pi up / U) does not depend on
    // the PRN: it is formed once per (Doppler bin, block) and shared by every
    // row of that bin.  The reduced form takes it for codes of N1 * 341
    // chips (its matrix-core phase transforms), with 16-bit code spectra.
    pl.ucode = 0;
    pl.boxcar = nullptr;
    if (allow_ucode && pl.L == pl.N) {
        const bool fp32 = pl.precision == UGSDR_PREC_FP32;
        bool ok = true;
        int64_t Lc = 0;
        for (int64_t p = 0; p < pl.P && ok; ++p) {
            const Replica& rep = ctx->replicas.find(r->replica_keys[p])->second;
            ok = rep.code_spec && (pl.Tsub > 1 || rep.repeats == 1) && pl.N % rep.code_len == 0 &&
                 (p == 0 || rep.code_len == Lc);
            Lc = rep.code_len;
        }
        if (ok && (fp32 ? upsampled_code_supported(Lc, pl.N / Lc) : reduced_ucode_supported(Lc, pl.N / Lc))) {
            pl.ucode = (int)(pl.N / Lc);
            for (int64_t p = 0; p < pl.P; ++p) {
                const Replica& rep = ctx->replicas.find(r->replica_keys[p])->second;
                pl.specs[p] = static_cast<const float2*>(rep.code_spec);
                if (!fp32) pl.spec_scale = rep.code_scale;
            }
            if (int rc = boxcar_table(ctx, pl.N, pl.ucode, &pl.boxcar, &pl.zbound)) return rc;
        }
    }
    // front end: unpack + mix + resample
    const int64_t fs_in = (int64_t)std::llround(r->fs_in), fs_acq = (int64_t)std::llround(r->fs_acq);
    pl.up = pl.down = 1;
    if (fs_in != fs_acq) {
        const int64_t g = gcd64(fs_acq, fs_in);
        const int64_t lcm = fs_acq / g * fs_in;
        pl.up = lcm / fs_in;
        pl.down = lcm / fs_acq;
    }
    pl.need = pl.N * pl.Tsub * pl.K;
    const int64_t avail = (pl.up == 1 && pl.down == 1) ? r->n_samples : (r->n_samples * pl.up + pl.down - 1) / pl.down;
    if (avail < pl.need) {
        set_error("signal shorter than n_blocks * n_fft at fs_acq");
        return UGSDR_ERR_SIZE;
    }
    pl.fmt = r->fmt;
    pl.n_in_used = std::min<int64_t>(r->n_samples, (pl.need * pl.down + pl.up - 1) / pl.up + 1);
    if (pl.fmt == UGSDR_FMT_BBP_DDC && pl.n_in_used >= ((int64_t)1 << 31)) {
        set_error("BBP_DDC requests take at most 2^31 - 1 samples");
        return UGSDR_ERR_SIZE;
    }
    pl.in_bytes = format_bytes(pl.fmt, pl.n_in_used, pl.spm_in, pl.epoch_bytes);
    double f = r->mix_frequency;
    if (f < 0) f = r->fs_in + f;  // Mixer::Translate folding (mixer.hpp:46-47)
    double ph = std::fmod(r->mix_phase, 2 * M_PI);
    if (ph < 0) ph += 2 * M_PI;
    pl.apply_tone = (f != 0.0 || ph != 0.0) ? 1 : 0;
    pl.f_turns = f / r->fs_in;
    pl.ph_turns = ph / (2 * M_PI);
    pl.ratio = r->fs_in / r->fs_acq;
    pl.intermediate_frequency = r->intermediate_frequency;
    pl.threshold = r->threshold;
    // Doppler classes
    pl.doppler.assign(r->doppler, r->doppler + pl.D);
    doppler_classes(r->doppler, pl.D, pl.N, r->fs_acq, pl.cls_frac, pl.cls, pl.shift);
    pl.F = (int)pl.cls_frac.size();
    pl.off_frac = 0;
    pl.off_dop = pl.off_frac + pl.cls_frac.size() * 8;
    pl.off_cls = pl.off_dop + pl.D * 8;
    pl.off_shift = pl.off_cls + pl.D * 4;
    pl.off_ucls = pl.off_shift + pl.D * 4;
    pl.off_ushift = pl.off_ucls + (pl.ucode ? pl.D * 4 : 0);
    pl.off_specs = (pl.off_ushift + (pl.ucode ? pl.D * 4 : 0) + 7) / 8 * 8;
    pl.table_bytes = pl.off_specs + pl.P * 8;
    return UGSDR_OK;
}

int enqueue_front(const AcqPlan& pl, const ugsdr_acq_request* r, AcqSlot& s, hipStream_t st, bool keep_surf,
                  FrontStep step) {
    if (step == FrontStep::Launch) return launch_front(pl, r, s, st);
    // tables
    s.host_tables.resize(pl.table_bytes);
    char* hp = s.host_tables.data();
    memcpy(hp + pl.off_frac, pl.cls_frac.data(), pl.cls_frac.size() * 8);
    memcpy(hp + pl.off_dop, pl.doppler.data(), pl.D * 8);
    memcpy(hp + pl.off_cls, pl.cls.data(), pl.D * 4);
    memcpy(hp + pl.off_shift, pl.shift.data(), pl.D * 4);
    if (pl.ucode) {
        int* uc = reinterpret_cast<int*>(hp + pl.off_ucls);
        int* us = reinterpret_cast<int*>(hp + pl.off_ushift);
        for (int64_t d = 0; d < pl.D; ++d) {
            uc[d] = (int)(d * pl.K);
            us[d] = 0;
        }
    }
    memcpy(hp + pl.off_specs, pl.specs.data(), pl.P * 8);
    UGSDR_HIP_TRY(s.tables.reserve(pl.table_bytes));
    UGSDR_HIP_TRY(hipMemcpyAsync(s.tables.ptr, hp, pl.table_bytes, hipMemcpyHostToDevice, st));
    // input
    const void* iq = r->iq;
    if (!r->iq_on_device) {
        if (pl.fmt == UGSDR_FMT_BBP_DDC)
            if (int rc = verify_bbp_headers(r->iq, pl.in_bytes / pl.epoch_bytes, pl.epoch_bytes)) return rc;
        UGSDR_HIP_TRY(s.iq.reserve(pl.in_bytes));
        UGSDR_HIP_TRY(hipMemcpyAsync(s.iq.ptr, r->iq, pl.in_bytes, hipMemcpyHostToDevice, st));
        iq = s.iq.ptr;
    }
    UGSDR_HIP_TRY(s.sig.reserve((size_t)pl.need * sizeof(float2)));
    UGSDR_HIP_TRY(s.rows.reserve((size_t)pl.P * pl.D * sizeof(RowStat)));
    if (keep_surf) UGSDR_HIP_TRY(s.surf.reserve((size_t)pl.P * pl.D * pl.L * sizeof(float)));
    (void)iq;
    if (step == FrontStep::Upload) return UGSDR_OK;
    return launch_front(pl, r, s, st);
}

// the front-end kernel alone (its tables and host IQ already uploaded)
int launch_front_args(const FrontArgs& fa, hipStream_t st) {
    const int blocks = (int)std::min<int64_t>((fa.n_out + 255) / 256, 4096);
    hipLaunchKernelGGL(front_kernel, dim3(blocks), dim3(256), 0, st, fa);
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

int launch_front(const AcqPlan& pl, const ugsdr_acq_request* r, AcqSlot& s, hipStream_t st) {
    FrontArgs fa{};
    fa.src.p = r->iq_on_device ? r->iq : s.iq.ptr;
    // decode only: the class transforms read the samples where they lie
    // (the plan holds n_in_used >= need samples)
    s.front_src = nullptr;
    if (pl.up == 1 && pl.down == 1 && !pl.apply_tone && (pl.fmt == UGSDR_FMT_CI16 || pl.fmt == UGSDR_FMT_CF32) &&
        pl.n_in_used >= pl.need) {
        s.front_src = fa.src.p;
        s.front_fmt = pl.fmt;
        return UGSDR_OK;
    }
    fa.src.fmt = pl.fmt;
    fa.src.spm = pl.spm_in;
    fa.src.epoch_bytes = pl.epoch_bytes;
    fa.n_in = pl.n_in_used;
    fa.n_out = pl.need;
    fa.apply_tone = pl.apply_tone;
    fa.f_turns = pl.f_turns;
    fa.ph_turns = pl.ph_turns;
    fa.up = pl.up;
    fa.down = pl.down;
    fa.out = s.sig.as<float2>();
    return launch_front_args(fa, st);
}

// Periodic fold: combined class spectrum of block k and class f from its T
// period sub-blocks, out = T sum_j e^(-j 2 pi j phi_f) sub_(kT+j, f).  With
// d L / fs = m + phi, sub-block j's tone restarts at e^(-j 2 pi d j L / fs) =
// e^(-j 2 pi j phi): the factor depends on the class only.
// Grid: chunks workgroups per (k, f), in one dimension.
__global__ void combine_periods_kernel(const float2* sub, float2* out, int64_t L, int F, int T, int chunks,
                                       const double* class_frac) {
    __shared__ float2 w[kMaxPeriods];
    const int kf = blockIdx.x / chunks, chunk = blockIdx.x - kf * chunks, k = kf / F, f = kf - k * F;
    for (int j = threadIdx.x; j < T; j += blockDim.x) {
        double t = (double)j * class_frac[f];
        t -= floor(t);
        double sn, cs;
        sincospi(2.0 * t, &sn, &cs);
        w[j] = make_float2((float)(T * cs), (float)(-T * sn));
    }
    __syncthreads();
    const float2* base = sub + ((int64_t)k * T * F + f) * L;
    for (int64_t i = chunk * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)chunks * blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        for (int j = 0; j < T; ++j) {
            const float2 x = base[(int64_t)j * F * L + i];
            acc.x = fmaf(w[j].x, x.x, fmaf(-w[j].y, x.y, acc.x));
            acc.y = fmaf(w[j].x, x.y, fmaf(w[j].y, x.x, acc.y));
        }
        out[(int64_t)kf * L + i] = acc;
    }
}

int enqueue_classes(const AcqPlan& pl, AcqSlot& s, hipStream_t st) {
    UGSDR_HIP_TRY(s.classes.reserve((size_t)pl.K * pl.F * pl.N * sizeof(float2)));
    FwdArgs a{};
    a.N = (int)pl.N;
    a.F = pl.F;
    a.sig = s.sig.as<float2>();
    if (s.front_src) {
        if (s.front_fmt == UGSDR_FMT_CI16)
            a.sig16 = static_cast<const short2*>(s.front_src);
        else
            a.sig = static_cast<const float2*>(s.front_src);
    }
    a.class_frac = reinterpret_cast<const double*>(s.tables.as<char>() + pl.off_frac);
    a.mode = 0;
    // the reduced form's Z is scaled per (bin, block) from the class blocks' max |X|
    auto zprep = [&]() -> int {
        if (!pl.ucode) return UGSDR_OK;
        if (pl.precision != UGSDR_PREC_FP32)
            if (int rc = enqueue_block_max(pl, s, st)) return rc;
        return launch_zprep(pl, s, st);
    };
    if (pl.Tsub == 1) {
        a.out = s.classes.as<float2>();
        if (int rc = launch_forward(a, (int)(pl.K * pl.F), st, s.fwd_scratch)) return rc;
        return zprep();
    }
    // every period sub-block's class spectra, then their combination per block
    UGSDR_HIP_TRY(s.classes_sub.reserve((size_t)pl.K * pl.Tsub * pl.F * pl.N * sizeof(float2)));
    a.out = s.classes_sub.as<float2>();
    if (int rc = launch_forward(a, (int)(pl.K * pl.Tsub * pl.F), st, s.fwd_scratch)) return rc;
    const int chunks = (int)std::min<int64_t>((pl.N + 255) / 256, 64);
    hipLaunchKernelGGL(combine_periods_kernel, dim3((unsigned)(pl.K * pl.F * chunks)), dim3(256), 0, st,
                       s.classes_sub.as<float2>(), s.classes.as<float2>(), pl.N, pl.F, (int)pl.Tsub, chunks,
                       a.class_frac);
    UGSDR_HIP_TRY(hipGetLastError());
    return zprep();
}

int default_threads(int64_t N) {
    if (const int t = corr1023_threads(N)) return t;
    if (N <= 8192) return 256;
    // four-step lengths: 1024 threads measured fastest on MI355X (BOC 4 ms,
    // N = 81840: 15.1 ms against 19.8 ms at 512)
    FourStep probe;
    if (N > kLdsCap && fourstep_ct((int)N, probe)) return 1024;
    return 512;
}

// The three-stage 1023*M kernel is identified by its own workgroup size
// (192 threads at N = 4092, 512 at 8184, 12276 and 16368, where it takes
// the place of the generic kernel of that size); the others select
// corr_small shapes.
std::vector<int> thread_candidates(int64_t N) {
    const int t1023 = corr1023_threads(N);
    switch (N) {
        case 4092: return {t1023, 256, 512};
        case 4200: return {256, 512};
        case 8184: return {t1023, 256};
        case 12276:
        case 16368: return {t1023, 1024};
        default: {
            FourStep probe;
            if (N > kLdsCap && fourstep_ct((int)N, probe)) return {512, 1024};
            return {default_threads(N)};
        }
    }
}

size_t corr_scratch_bytes(const AcqPlan& p0, int cus) {
    if (p0.precision != UGSDR_PREC_FP32) return p0.ucode ? 0 : reduced_scratch_bytes(p0.N, (int)p0.K, cus);
    if (p0.ucode && p0.N / p0.ucode != 1023) return 0;   // phase transforms: no scratch
    if (!p0.small) return (2 * (size_t)p0.N + (p0.L + 1) / 2) * sizeof(float2) * (size_t)cus;
    if (corr1023_supports(p0.N, p0.L)) return (size_t)3 * corr1023_grid(p0.N, cus, p0.ucode != 0) * p0.L * 4;
    return 0;
}

int enqueue_corr(const std::vector<const AcqPlan*>& plans, const std::vector<AcqSlot*>& slots, hipStream_t st,
                 DevBuf& scratch, int cus, int threads) {
    if (plans.empty() || plans.size() > (size_t)kMaxBatch || plans.size() != slots.size()) {
        set_error("a correlation batch holds 1 to " + std::to_string(kMaxBatch) + " requests");
        return UGSDR_ERR_ARG;
    }
    const AcqPlan& p0 = *plans[0];
    CorrArgs c{};
    ReqDesc* reqs = c.reqs;
    int rows = 0, units = 0, maxK = 1;
    for (size_t i = 0; i < plans.size(); ++i) {
        const AcqPlan& pl = *plans[i];
        AcqSlot& s = *slots[i];
        if (pl.bucket() != p0.bucket() || pl.L != p0.L) {
            set_error("batch mixes transform shapes");
            return UGSDR_ERR_ARG;
        }
        char* dp = s.tables.as<char>();
        ReqDesc& q = reqs[i];
        q.X = pl.ucode ? s.zbuf.as<float2>() : s.classes.as<float2>();
        q.specs = reinterpret_cast<const float2* const*>(dp + pl.off_specs);
        q.cls = reinterpret_cast<const int*>(dp + (pl.ucode ? pl.off_ucls : pl.off_cls));
        q.shift = reinterpret_cast<const int*>(dp + (pl.ucode ? pl.off_ushift : pl.off_shift));
        q.rowstats = s.rows.as<RowStat>();
        q.surfaces = s.surf.ptr ? s.surf.as<float>() : nullptr;
        q.K = (int)pl.K;
        q.F = pl.ucode ? 1 : pl.F;
        q.D = (int)pl.D;
        q.P = (int)pl.P;
        q.row_start = rows;
        rows += (int)(pl.P * pl.D);
        maxK = std::max(maxK, (int)pl.K);
    }
    const int T = threads > 0 ? threads : default_threads(p0.N);
    // the upsampled-code search runs on its own kernels whatever variant the
    // length is tuned to: the wave-autonomous ones for 1023-chip codes (the
    // three-stage launch below), the phase transforms for longer codes
    const bool phase_tf = p0.ucode && p0.N / p0.ucode != 1023;
    const bool three_stage =
        !phase_tf && p0.small && (p0.ucode || T == corr1023_threads(p0.N)) && corr1023_supports(p0.N, p0.L);
    if (p0.ucode && !phase_tf && !three_stage) {
        set_error("upsampled-code search without a three-stage kernel for its length");
        return UGSDR_ERR_UNSUPPORTED;
    }
    // three-stage kernel: whole rounds of rows one per workgroup, the rest
    // (when rows have several blocks) split block-wise as the batch tail
    int full_rows = rows;
    if (three_stage && maxK > 1) {
        const int G = corr1023_grid(p0.N, cus, p0.ucode != 0);
        full_rows = rows / G * G;
    }
    for (size_t i = 0; i < plans.size(); ++i) {
        ReqDesc& q = reqs[i];
        const int n = q.P * q.D;
        const int a0 = std::max(q.row_start, full_rows) - q.row_start;
        q.tail_row0 = std::min(a0, n);
        q.tail_unit_start = units;
        units += (n - q.tail_row0) * q.K;
    }
    c.plan = p0.plan;
    c.fs = p0.fs;
    c.N = (int)p0.N;
    c.L = (int)p0.L;
    c.fold_mode = p0.fold_mode;
    c.rows = rows;
    c.n_req = (int)plans.size();
    c.inv_n = (float)(1.0 / (double)p0.N);
    c.full_rows = full_rows;
    c.units = units;
    c.ucode = p0.ucode;
    int rc = UGSDR_OK;
    if (phase_tf) {
        rc = launch_corr1023p(c, cus, st);
    } else if (three_stage) {
        rc = launch_corr1023(c, cus, scratch, st);
    } else if (p0.small) {
        const int grid = rows;
        const int Ts = T == corr1023_threads(c.N) ? (c.N <= 8192 ? 256 : 512) : T;
        switch (c.N) {
            case 4092:
                rc = Ts == 512 ? launch_corr_small<512, 4096, P4092>(c, grid, st)
                              : launch_corr_small<256, 4096, P4092>(c, grid, st);
                break;
            case 8184:
                rc = Ts == 512 ? launch_corr_small<512, 8192, P8184>(c, grid, st)
                              : launch_corr_small<256, 8192, P8184>(c, grid, st);
                break;
            case 12276:
                rc = Ts == 1024 ? launch_corr_small<1024, kLdsCap, P12276>(c, grid, st)
                               : launch_corr_small<512, kLdsCap, P12276>(c, grid, st);
                break;
            case 16368:
                rc = Ts == 1024 ? launch_corr_small<1024, kLdsCap, P16368>(c, grid, st)
                               : launch_corr_small<512, kLdsCap, P16368>(c, grid, st);
                break;
            case 4200:
                rc = Ts == 512 ? launch_corr_small<512, 8192, P4200>(c, grid, st)
                              : launch_corr_small<256, 8192, P4200>(c, grid, st);
                break;
            default:
                if (c.N <= 4096) rc = launch_corr_small<256, 4096, PlanRT>(c, grid, st);
                else if (c.N <= 8192) rc = launch_corr_small<256, 8192, PlanRT>(c, grid, st);
                else rc = launch_corr_small<512, kLdsCap, PlanRT>(c, grid, st);
        }
    } else {
        const int grid = std::min(rows, cus);
        const size_t per_wg = (2 * (size_t)c.N + (c.L + 1) / 2) * sizeof(float2);
        UGSDR_HIP_TRY(scratch.reserve(per_wg * grid));
        c.scratch = scratch.as<float2>();
        FourStep probe;
        const bool big = T == 1024;
        if (!fourstep_ct(c.N, probe)) rc = launch_corr_large<512, PlanRT, PlanRT>(c, grid, st);
        else if (c.fs.n2 == 175)
            rc = big ? launch_corr_large<1024, P152, P175>(c, grid, st) : launch_corr_large<512, P152, P175>(c, grid, st);
        else if (c.fs.n2 == 256) {
            if (c.fs.n1 == 128)
                rc = big ? launch_corr_large<1024, P128, P256>(c, grid, st)
                         : launch_corr_large<512, P128, P256>(c, grid, st);
            else
                rc = big ? launch_corr_large<1024, P256, P256>(c, grid, st)
                         : launch_corr_large<512, P256, P256>(c, grid, st);
        } else switch (c.fs.n1) {
            case 60: rc = big ? launch_corr_large<1024, P60, P341>(c, grid, st)
                              : launch_corr_large<512, P60, P341>(c, grid, st); break;
            case 112: rc = big ? launch_corr_large<1024, P112, P341>(c, grid, st)
                               : launch_corr_large<512, P112, P341>(c, grid, st); break;
            case 120: rc = big ? launch_corr_large<1024, P120, P341>(c, grid, st)
                               : launch_corr_large<512, P120, P341>(c, grid, st); break;
            case 180: rc = big ? launch_corr_large<1024, P180, P341>(c, grid, st)
                               : launch_corr_large<512, P180, P341>(c, grid, st); break;
            default: rc = big ? launch_corr_large<1024, P240, P341>(c, grid, st)
                              : launch_corr_large<512, P240, P341>(c, grid, st); break;
        }
    }
    if (rc) return rc;
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

int presize_slot(const AcqPlan& pl, AcqSlot& s, bool peaks) {
    const size_t kfn = (size_t)pl.K * pl.F * pl.N;
    UGSDR_HIP_TRY(s.classes.reserve(kfn * sizeof(float2)));
    if (pl.Tsub > 1) UGSDR_HIP_TRY(s.classes_sub.reserve(kfn * pl.Tsub * sizeof(float2)));
    if (pl.fwd_scratch_needed()) UGSDR_HIP_TRY(s.fwd_scratch.reserve(kfn * pl.Tsub * sizeof(float2)));
    const bool fp32 = pl.precision == UGSDR_PREC_FP32;
    if (pl.ucode) UGSDR_HIP_TRY(s.zbuf.reserve((size_t)pl.D * pl.K * pl.N * (fp32 ? 8 : 4)));
    if (!fp32) {
        if (!pl.ucode) UGSDR_HIP_TRY(s.classes_r.reserve(kfn * 4));
        UGSDR_HIP_TRY(s.block_scale.reserve((size_t)pl.K * pl.F * 8 + (pl.ucode ? (size_t)pl.D * pl.K * 4 : 0)));
    }
    UGSDR_HIP_TRY(s.results.reserve((size_t)pl.P * sizeof(ugsdr_acq_result) + (peaks ? (size_t)pl.P * pl.L * 4 : 0)));
    return UGSDR_OK;
}

int enqueue_final(const AcqPlan& pl, AcqSlot& s, hipStream_t st, bool peaks) {
    UGSDR_HIP_TRY(s.results.reserve((size_t)pl.P * sizeof(ugsdr_acq_result) + (peaks ? (size_t)pl.P * pl.L * 4 : 0)));
    FinalArgs fa{};
    fa.rowstats = s.rows.as<RowStat>();
    fa.doppler = reinterpret_cast<const double*>(s.tables.as<char>() + pl.off_dop);
    fa.D = (int)pl.D;
    fa.P = (int)pl.P;
    fa.ratio = pl.ratio;
    fa.intermediate_frequency = pl.intermediate_frequency;
    fa.threshold = pl.threshold;
    fa.reshape = pl.reshape;
    fa.out = s.results.as<ugsdr_acq_result>();
    fa.surfaces = s.surf.ptr ? s.surf.as<float>() : nullptr;
    fa.peaks = peaks ? reinterpret_cast<float*>(s.results.as<char>() + (size_t)pl.P * sizeof(ugsdr_acq_result))
                     : nullptr;
    fa.L = (int)pl.L;
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)pl.P), dim3(256), 0, st, fa);
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

// ------------------------------------------------------------------ autotuner
int autotune_corr(ugsdr_hip_ctx* ctx, int64_t N, int* best, std::vector<float>* ms_out) {
    const std::vector<int> cands = thread_candidates(N);
    if (ms_out) ms_out->assign(cands.size(), 0.f);
    if (cands.size() == 1) {
        ctx->tuned[N] = cands[0];
        if (best) *best = cands[0];
        return UGSDR_OK;
    }
    // synthetic batch that fills the device: 32 PRNs x 32 Doppler bins, 1
    // block; the four-step lengths take 4 blocks, since how their variants
    // rank depends on the blocks a row accumulates (N = 81840: 512 threads win
    // a 1-block batch, 1024 a 5-block search by 1.5x)
    const int P = 32, D = 32;
    const int K = N > kLdsCap ? 4 : 1;
    uint64_t x = 0x9E3779B97F4A7C15ull ^ (uint64_t)N;
    auto next = [&x]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    std::vector<float> iq(2 * (size_t)N * K);
    for (auto& v : iq) v = (float)((int)(next() % 255) - 127);
    std::vector<uint64_t> keys;
    std::vector<int8_t> chips(1023);
    int rc = UGSDR_OK;
    for (int p = 0; p < P && !rc; ++p) {
        for (auto& c : chips) c = (next() >> 63) ? 1 : -1;
        Replica rep;
        rep.n_fft = N;
        if (hipMalloc(&rep.spec, (size_t)N * sizeof(float2)) != hipSuccess) {
            set_error("autotune: out of device memory");
            rc = UGSDR_ERR_NOMEM;
            break;
        }
        rc = replica_spectrum(ctx, chips.data(), 1023, 1, nullptr, N, static_cast<float2*>(rep.spec));
        const uint64_t key = ~(uint64_t)p;  // keys from the top of the space, removed below
        ctx->replicas[key] = rep;
        keys.push_back(key);
    }
    std::vector<double> dop(D);
    const double fs = (double)N * 1000.0;
    for (int d = 0; d < D; ++d) dop[d] = -4000.0 + 250.0 * d;
    ugsdr_acq_request r{};
    r.iq = iq.data();
    r.fmt = UGSDR_FMT_CF32;
    r.n_samples = N * K;
    r.fs_in = r.fs_acq = fs;
    r.n_fft = N;
    r.n_blocks = K;
    r.surface_len = N;
    r.reshape = 1;
    r.coherent = 1;
    r.doppler = dop.data();
    r.n_doppler = D;
    r.replica_keys = keys.data();
    r.n_prn = P;
    r.threshold = 3.3;
    AcqPlan pl;
    AcqSlot slot;
    DevBuf scratch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = make_acq_plan(ctx, &r, pl);
    if (!rc) rc = enqueue_front(pl, &r, slot, ctx->stream, false);
    if (!rc) rc = enqueue_classes(pl, slot, ctx->stream);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        set_error("autotune: event creation failed");
        rc = UGSDR_ERR_HIP;
    }
    int chosen = cands[0];
    float best_ms = 1e30f;
    const int cus = device_cu_count(ctx->device);
    for (size_t i = 0; i < cands.size() && !rc; ++i) {
        rc = enqueue_corr({&pl}, {&slot}, ctx->stream, scratch, cus, cands[i]);  // warm-up
        if (rc) break;
        (void)hipEventRecord(e0, ctx->stream);
        for (int it = 0; it < 3 && !rc; ++it) rc = enqueue_corr({&pl}, {&slot}, ctx->stream, scratch, cus, cands[i]);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            set_error("autotune: kernel failed");
            rc = UGSDR_ERR_HIP;
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= 3.f;
        if (ms_out) (*ms_out)[i] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            chosen = cands[i];
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (uint64_t k : keys) {
        (void)hipFree(ctx->replicas[k].spec);
        ctx->replicas.erase(k);
    }
    slot.release_all();
    scratch.release();
    if (rc) return rc;
    ctx->tuned[N] = chosen;
    if (best) *best = chosen;
    return UGSDR_OK;
}

int tuned_threads(ugsdr_hip_ctx* ctx, int64_t N) {
    auto it = ctx->tuned.find(N);
    if (it != ctx->tuned.end()) return it->second;
    if (!ctx->autotune || thread_candidates(N).size() == 1) return default_threads(N);
    int best = 0;
    if (autotune_corr(ctx, N, &best, nullptr) != UGSDR_OK) {
        ctx->tuned[N] = default_threads(N);  // keep serving with the default shape
        return default_threads(N);
    }
    return best;
}

// Key of a request's device work: everything but the host IQ pointer (host
// IQ is uploaded on every call, outside any graph).
static std::string graph_key(const ugsdr_acq_request* r, bool surfaces, bool peaks, int threads) {
    ugsdr_acq_request c = *r;
    if (!c.iq_on_device) c.iq = nullptr;
    c.doppler = nullptr;
    c.replica_keys = nullptr;
    c.hypotheses = 0;
    std::string k(reinterpret_cast<const char*>(&c), sizeof c);
    k.append(reinterpret_cast<const char*>(r->doppler), (size_t)r->n_doppler * sizeof(double));
    k.append(reinterpret_cast<const char*>(r->replica_keys), (size_t)r->n_prn * sizeof(uint64_t));
    const int32_t tail[3] = {surfaces ? 1 : 0, peaks ? 1 : 0, threads};
    k.append(reinterpret_cast<const char*>(tail), sizeof tail);
    return k;
}

// front end, class FFTs, correlation and selection of one request on st
static int enqueue_search(ugsdr_hip_ctx* ctx, const AcqPlan& pl, const ugsdr_acq_request* r, AcqSlot& s,
                          hipStream_t st, bool keep_surf, bool peaks, int threads, FrontStep front) {
    ctx->mark(0);
    if (int rc = enqueue_front(pl, r, s, st, keep_surf, front)) return rc;
    ctx->mark(1);
    if (int rc = enqueue_classes(pl, s, st)) return rc;
    ctx->mark(2);
    if (int rc = enqueue_corr({&pl}, {&s}, st, ctx->scratch, device_cu_count(ctx->device), threads)) return rc;
    ctx->mark(3);
    if (int rc = enqueue_final(pl, s, st, peaks)) return rc;
    ctx->mark(4);
    return UGSDR_OK;
}

int acquire_impl(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, ugsdr_acq_result* out, float* surfaces,
                 float* peaks) {
    UGSDR_CHECK_ARG(out, "null result");
    if (r && r->precision != UGSDR_PREC_FP32) return acquire_reduced(ctx, r, out, surfaces, peaks);
    AcqPlan pl;
    if (int rc = make_acq_plan(ctx, r, pl)) return rc;
    ctx->last_n = pl.N;
    ctx->last_periods = (int)pl.Tsub;
    ctx->last_ucode = pl.ucode;
    AcqSlot& s = ctx->slot;
    const bool keep_surf = surfaces || peaks;
    hipStream_t st = ctx->stream;
    // the upsampled-code search runs its own kernels whatever the length's
    // variant: nothing to tune
    const int threads = pl.ucode ? 0 : tuned_threads(ctx, pl.N);
    // A request shape seen twice in a row is captured into a hipGraph on the
    // second call and replayed afterwards: the kernels and their arguments
    // (device buffers included) are the same, only the host IQ upload and
    // the result copies run outside the graph.  Any other call drops the
    // graph, since it may move the buffers the graph refers to.
    auto& gc = ctx->graph;
    const bool graphs = gc.enabled && !ctx->profile;
    const std::string key = graphs ? graph_key(r, surfaces != nullptr, peaks != nullptr, threads) : std::string();
    if (graphs && gc.exec && key == gc.key) {
        if (!r->iq_on_device) {
            if (pl.fmt == UGSDR_FMT_BBP_DDC)
                if (int rc = verify_bbp_headers(r->iq, pl.in_bytes / pl.epoch_bytes, pl.epoch_bytes)) return rc;
            UGSDR_HIP_TRY(hipMemcpyAsync(s.iq.ptr, r->iq, pl.in_bytes, hipMemcpyHostToDevice, st));
        }
        UGSDR_HIP_TRY(hipGraphLaunch(gc.exec, st));
        ++gc.replays;
    } else if (graphs && key == gc.pending) {
        gc.drop();
        if (int rc = enqueue_front(pl, r, s, st, keep_surf, FrontStep::Upload)) return rc;
        UGSDR_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        int rc = enqueue_search(ctx, pl, r, s, st, keep_surf, peaks != nullptr, threads, FrontStep::Launch);
        g_capturing = false;
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        hipGraphExec_t exec = nullptr;
        if (!rc && ec == hipSuccess && g && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess) {
            gc.exec = exec;
            gc.key = key;
            gc.pending.clear();
        } else {
            gc.enabled = false;  // this context runs eagerly from now on
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        if (gc.exec) {
            UGSDR_HIP_TRY(hipGraphLaunch(gc.exec, st));
        } else if (int rc2 = enqueue_search(ctx, pl, r, s, st, keep_surf, peaks != nullptr, threads,
                                            FrontStep::Launch)) {
            return rc2;
        }
    } else {
        gc.drop();
        if (!keep_surf) s.surf.release();
        if (int rc = enqueue_search(ctx, pl, r, s, st, keep_surf, peaks != nullptr, threads, FrontStep::All))
            return rc;
        if (graphs) gc.pending = key;
    }
    // results and peaks lie together in s.results: one copy to page-locked
    // staging, then to the caller's memory
    const size_t res_bytes = (size_t)pl.P * sizeof(ugsdr_acq_result);
    const size_t peak_bytes = peaks ? (size_t)pl.P * pl.L * 4 : 0;
    UGSDR_HIP_TRY(ctx->results_host.reserve(res_bytes + peak_bytes));
    UGSDR_HIP_TRY(hipMemcpyAsync(ctx->results_host.ptr, s.results.ptr, res_bytes + peak_bytes, hipMemcpyDeviceToHost,
                                 st));
    if (surfaces)
        UGSDR_HIP_TRY(hipMemcpyAsync(surfaces, s.surf.ptr, (size_t)pl.P * pl.D * pl.L * 4, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    memcpy(out, ctx->results_host.ptr, res_bytes);
    if (peaks) memcpy(peaks, static_cast<const char*>(ctx->results_host.ptr) + res_bytes, peak_bytes);
    if (ctx->profile)
        for (int i = 0; i < 4; ++i) UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[i], ctx->ev[i], ctx->ev[i + 1]));
    return UGSDR_OK;
}

void reduce_hypotheses(const ugsdr_acq_result* in, int64_t n, int h, double threshold, ugsdr_acq_result* out) {
    for (int64_t g = 0; g < n / h; ++g) {
        const ugsdr_acq_result* run = in + g * h;
        int best = 0;
        for (int k = 1; k < h; ++k)
            if (run[best].level < run[k].level) best = k;
        out[g] = run[best];
        out[g].detected = run[best].snr > threshold ? 1 : 0;
    }
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
    if (!req || req->hypotheses <= 1 || !out) return ugsdr_hip::acquire_impl(ctx, req, out, surfaces, peaks);
    if (req->n_prn <= 0 || req->n_prn % req->hypotheses) {
        ugsdr_hip::set_error("n_prn must be a multiple of the hypothesis count");
        return UGSDR_ERR_ARG;
    }
    std::vector<ugsdr_acq_result> all((size_t)req->n_prn);
    if (int rc = ugsdr_hip::acquire_impl(ctx, req, all.data(), surfaces, peaks)) return rc;
    ugsdr_hip::reduce_hypotheses(all.data(), req->n_prn, req->hypotheses, req->threshold, out);
    return UGSDR_OK;
}

extern "C" int ugsdr_hip_autotune(ugsdr_hip_ctx* ctx, int64_t n_fft, int32_t* threads, float* ms_out, int32_t n_ms) {
    if (!ctx || n_fft <= 0) {
        ugsdr_hip::set_error("bad arguments");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    int best = 0;
    std::vector<float> ms;
    if (int rc = ugsdr_hip::autotune_corr(ctx, n_fft, &best, &ms)) return rc;
    if (threads) *threads = best;
    for (int i = 0; ms_out && i < n_ms && i < (int)ms.size(); ++i) ms_out[i] = ms[i];
    return UGSDR_OK;
}

extern "C" int ugsdr_hip_corr_variants(int64_t n_fft, int32_t* threads_out, int32_t n) {
    if (n_fft <= 0 || (n > 0 && !threads_out)) {
        ugsdr_hip::set_error("bad arguments");
        return UGSDR_ERR_ARG;
    }
    const std::vector<int> c = ugsdr_hip::thread_candidates(n_fft);
    for (int i = 0; i < n && i < (int)c.size(); ++i) threads_out[i] = c[i];
    return (int)c.size();
}

extern "C" int ugsdr_hip_set_threads(ugsdr_hip_ctx* ctx, int64_t n_fft, int32_t threads) {
    if (!ctx || n_fft <= 0) {
        ugsdr_hip::set_error("bad arguments");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (threads == 0) {
        ctx->tuned.erase(n_fft);
        return UGSDR_OK;
    }
    const std::vector<int> c = ugsdr_hip::thread_candidates(n_fft);
    if (std::find(c.begin(), c.end(), (int)threads) == c.end()) {
        ugsdr_hip::set_error("no correlation variant with " + std::to_string(threads) + " threads for length " +
                             std::to_string(n_fft));
        return UGSDR_ERR_ARG;
    }
    ctx->tuned[n_fft] = threads;
    return UGSDR_OK;
}

extern "C" int ugsdr_hip_set_graphs(ugsdr_hip_ctx* ctx, int enable) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->graph.drop();
    ctx->graph.pending.clear();
    ctx->graph.enabled = enable != 0;
    return UGSDR_OK;
}

extern "C" int64_t ugsdr_hip_graph_replays(ugsdr_hip_ctx* ctx) { return ctx ? ctx->graph.replays : -1; }

extern "C" int ugsdr_hip_last_search_form(ugsdr_hip_ctx* ctx, int64_t* n_transform, int32_t* periods,
                                          int32_t* samples_per_chip) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (n_transform) *n_transform = ctx->last_n;
    if (periods) *periods = ctx->last_periods;
    if (samples_per_chip) *samples_per_chip = ctx->last_ucode;
    return UGSDR_OK;
}

extern "C" int ugsdr_hip_set_autotune(ugsdr_hip_ctx* ctx, int enable) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->autotune = enable != 0;
    return UGSDR_OK;
}
