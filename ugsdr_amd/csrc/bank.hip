// Tracking correlator bank: many channels x taps over consecutive 1 ms blocks.
//
// NCO mode: per sample, FP64 carrier phase and code phase, FP32 sincos of the
// reduced carrier phase, replica chip looked up from the channel's code in
// LDS, FP32 I/Q accumulators, workgroup reduction per tap.
// Ref-split mode: Tracker::GetEpl semantics (in-place Mixer::Translate, then
// TrackingParameters::CorrelateSplit against a 3-period upsampled replica).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

constexpr int kBankT = 256;
constexpr int kMaxTaps = 8;

struct SplitTap {
    int first_len;
    int first_phase;
    int second_phase;
    float pad;
    double rel;  // fmod(phase, spm) / spm
};

struct BankArgs {
    const void* iq;
    int fmt;
    int64_t spm;
    int n_blocks;
    int n_ch;
    int n_taps;
    double fs;
    // nco mode
    const double* ch;            // n_ch x 4: f, theta0 (cycles), rate, tau0 (chips)
    const signed char* codes;    // n_ch x code_len
    int64_t code_len;
    double taps[kMaxTaps];       // chips
    int tap_margin;              // nco mode: chips of wrap padding on each side of the LDS code table
    int q_entries;               // quarter-table kernel: table entries (a multiple of 16)
    const struct QSeed* q_seed;  // quarter-table kernel: n_ch x n_blocks per-block seeds
    // ref_split mode
    const float* code3;          // n_ch x code3_len
    int64_t code3_len;
    const SplitTap* split;       // n_ch x n_taps
    const double* carrier;       // n_ch x 2: f_carr, phase0 (rad)
    float2* out;                 // n_blocks x n_ch x n_taps
};

__device__ __forceinline__ float2 bank_load(const void* p, int fmt, int64_t i) {
    if (fmt == UGSDR_FMT_CI16) {
        const short2 v = static_cast<const short2*>(p)[i];
        return make_float2((float)v.x, (float)v.y);
    }
    if (fmt == UGSDR_FMT_CI8) {
        const char2 v = static_cast<const char2*>(p)[i];
        return make_float2((float)v.x, (float)v.y);
    }
    return static_cast<const float2*>(p)[i];
}

template <int NT>
__device__ __forceinline__ void bank_reduce_store(float2 (&acc)[NT], int n_taps, float2* out, float2* red) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        if (t >= n_taps) break;
        acc[t].x = wave_sum(acc[t].x);
        acc[t].y = wave_sum(acc[t].y);
        if (l == 0) red[w * NT + t] = acc[t];
    }
    __syncthreads();
    if (threadIdx.x < n_taps) {
        float2 s = make_float2(0.f, 0.f);
        for (int i = 0; i < kBankT / 64; ++i) s = cadd(s, red[i * NT + threadIdx.x]);
        out[threadIdx.x] = s;
    }
}

// One workgroup per (channel, block).  The code sits in LDS with `margin`
// chips of wrap padding on each side, so a tap's chip is base + tap's
// integer part + (frac(u) >= the tap's threshold), with no modulo per tap:
// floor(u + tap) = floor(u) + floor(tap) + [frac(u) >= 1 - frac(tap)].  The
// carrier is an FP64 phasor per thread advanced by e^(j 2 pi T df) per
// sample slot (the same rate as the FP32 vector unit on this GPU), started
// from the thread's own FP64 phase.
__global__ __launch_bounds__(kBankT) void bank_nco_kernel(BankArgs a) {
    const int c = blockIdx.x, b = blockIdx.y;  // the channels of one block run together (IQ block reused from L2)
    extern __shared__ __align__(16) unsigned char smem[];
    signed char* code = reinterpret_cast<signed char*>(smem);
    const int Li = (int)a.code_len, M = a.tap_margin;
    float2* red = reinterpret_cast<float2*>(smem + ((Li + 2 * M + 15) / 16) * 16);
    for (int i = threadIdx.x; i < Li + 2 * M; i += kBankT) {
        int j = i - M;
        j += j < 0 ? Li : 0;
        j -= j >= Li ? Li : 0;
        code[i] = a.codes[(int64_t)c * a.code_len + j];
    }
    const double f = a.ch[4 * c], th0 = a.ch[4 * c + 1], rate = a.ch[4 * c + 2], tau0 = a.ch[4 * c + 3];
    const double L = (double)a.code_len;
    const double n0 = (double)b * (double)a.spm;
    double theta = fma(f, n0 / a.fs, th0);
    theta -= floor(theta);
    double tau = fma(rate, n0 / a.fs, tau0);
    tau = fmod(tau, L);
    const double df = f / a.fs, dr = rate / a.fs, invL = 1.0 / L;
    int ti[kMaxTaps];
    double thr[kMaxTaps];
#pragma unroll
    for (int t = 0; t < kMaxTaps; ++t) {
        const double tp = t < a.n_taps ? a.taps[t] : 0.0, fl = floor(tp);
        ti[t] = (int)fl + M;
        thr[t] = 1.0 - (tp - fl);
    }
    // e^(-j 2 pi ph) for this thread's first sample and its step per slot
    double zc, zs, wc, ws;
    {
        double ph = fma((double)threadIdx.x, df, theta);
        ph -= rint(ph);
        sincospi(2.0 * ph, &zs, &zc);
        zs = -zs;
        double st = (double)kBankT * df;
        st -= rint(st);
        sincospi(2.0 * st, &ws, &wc);
        ws = -ws;
    }
    float2 acc[kMaxTaps];
#pragma unroll
    for (int t = 0; t < kMaxTaps; ++t) acc[t] = make_float2(0.f, 0.f);
    __syncthreads();
    const void* blk = a.fmt == UGSDR_FMT_CI16 ? (const void*)(static_cast<const short2*>(a.iq) + b * a.spm)
                      : a.fmt == UGSDR_FMT_CI8 ? (const void*)(static_cast<const char2*>(a.iq) + b * a.spm)
                                               : (const void*)(static_cast<const float2*>(a.iq) + b * a.spm);
    for (int64_t n = threadIdx.x; n < a.spm; n += kBankT) {
        const float2 x = bank_load(blk, a.fmt, n);
        const float co = (float)zc, s = (float)zs;
        // x * exp(-j 2 pi ph)
        const float2 w = make_float2(fmaf(x.x, co, -x.y * s), fmaf(x.y, co, x.x * s));
        const double nzc = fma(zc, wc, -zs * ws);
        zs = fma(zc, ws, zs * wc);
        zc = nzc;
        // code phase in chips (FP64); floor(u) mod L once per sample
        const double u = fma((double)n, dr, tau);
        const double fu = floor(u);
        const double frac = u - fu;
        int base = (int)fma(-L, floor(fu * invL), fu);
        base -= base >= Li ? Li : 0;
        base += base < 0 ? Li : 0;
#pragma unroll
        for (int t = 0; t < kMaxTaps; ++t) {
            if (t >= a.n_taps) break;
            const int idx = base + ti[t] + (frac >= thr[t] ? 1 : 0);
            const float cv = (float)code[idx];
            acc[t].x = fmaf(w.x, cv, acc[t].x);
            acc[t].y = fmaf(w.y, cv, acc[t].y);
        }
    }
    bank_reduce_store<kMaxTaps>(acc, a.n_taps, a.out + ((int64_t)b * a.n_ch + c) * a.n_taps, red);
}

// NCO mode, several channels per workgroup and several blocks per workgroup.
//
// Each thread takes samples n = tid + s T of a block (T = 256, s the slot),
// loads each sample once for the workgroup's CH channels and, per channel:
//   * mixes it with an FP32 carrier phasor z = e^(-j 2 pi phase(n)), which
//     advances by one complex multiply with e^(-j 2 pi T df) per slot and is
//     re-seeded every kReseed slots from the FP64 phase (sincospif of the
//     reduced phase), so the recurrence error stays below
//     kReseed (2 + sqrt 2) 2^-24 (about 3e-6 at kReseed = 32);
//   * keeps the code phase as a 20.44 fixed-point count of chips, started from
//     the FP64 code phase of the thread's first sample and advanced by the
//     integer increment round(T dr 2^44) per slot (drift below 2^-45 chip per
//     slot, the FP64 arithmetic's own order), so a tap's chip index is the
//     integer part, the tap's integer offset and the carry of the 32-bit
//     fraction plus the tap's fraction; one conditional subtraction of L
//     (min(b, b - L) on unsigned values) keeps the index within the code;
//   * accumulates each tap in FP32.
// The channels' codes sit in LDS as int8 with `margin` chips of wrap padding
// on each side; the FP64 phases of each block start are computed per thread.
constexpr int kBankCH = 4;    // channels per workgroup
constexpr int kBankBB = 4;    // blocks per workgroup
constexpr int kReseed = 32;   // slots between carrier re-seeds
constexpr int kQSlots = 256;  // quarter-table kernel: most slots per block (spm <= kQSlots T)
constexpr int kQWrap = 8;     // quarter-table kernel: slots between code-index wrap checks
constexpr int kQSub = 1;      // quarter-table kernel: sub-groups of kBankT threads per workgroup
constexpr double kFix = 17592186044416.0;   // 2^44

template <int NT>
__global__ __launch_bounds__(kBankT) void bank_nco_multi(BankArgs a) {
    constexpr int CH = kBankCH;
    extern __shared__ __align__(16) unsigned char smem[];
    const int Li = (int)a.code_len, M = a.tap_margin, LT = Li + 2 * M;
    const int c0 = blockIdx.x * CH;
    const int nch = min(CH, a.n_ch - c0);
    signed char* code = reinterpret_cast<signed char*>(smem);
    float2* red = reinterpret_cast<float2*>(smem + ((CH * LT + 15) / 16) * 16);
    for (int i = threadIdx.x; i < CH * LT; i += kBankT) {
        const int c = i / LT, k = i - c * LT;
        int j = k - M;
        j += j < 0 ? Li : 0;
        j -= j >= Li ? Li : 0;
        code[i] = c < nch ? a.codes[(int64_t)(c0 + c) * a.code_len + j] : (signed char)1;
    }
    // taps: integer part (plus the margin) and 32-bit fraction
    int ti[NT];
    uint32_t tf[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double tp = t < a.n_taps ? a.taps[t] : 0.0, fl = floor(tp);
        ti[t] = (int)fl + M;
        tf[t] = (uint32_t)(uint64_t)llrint((tp - fl) * 4294967296.0);
    }
    double f[CH], th0[CH], rate[CH], tau0[CH];
    float2 w[CH];
    uint64_t inc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int cc = c0 + min(c, nch - 1);
        f[c] = a.ch[4 * cc];
        th0[c] = a.ch[4 * cc + 1];
        rate[c] = a.ch[4 * cc + 2];
        tau0[c] = a.ch[4 * cc + 3];
        double st = (double)kBankT * (f[c] / a.fs);
        st -= rint(st);
        double s, co;
        sincospi(2.0 * st, &s, &co);
        w[c] = make_float2((float)co, (float)-s);
        inc[c] = (uint64_t)llrint((double)kBankT * (rate[c] / a.fs) * kFix);
    }
    const double L = (double)a.code_len;
    const uint32_t Lu = (uint32_t)Li;
    __syncthreads();
    const int b_end = min(a.n_blocks, (int)(blockIdx.y + 1) * kBankBB);
    for (int b = blockIdx.y * kBankBB; b < b_end; ++b) {
        const double n0 = (double)b * (double)a.spm;
        double theta[CH];
        float2 z[CH];
        uint64_t u[CH];
        float2 acc[CH][NT];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            double th = fma(f[c], n0 / a.fs, th0[c]);
            theta[c] = th - floor(th);
            double tau = fmod(fma(rate[c], n0 / a.fs, tau0[c]), L);
            tau += tau < 0 ? L : 0.0;
            u[c] = (uint64_t)((tau + (double)threadIdx.x * (rate[c] / a.fs)) * kFix);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[c][t] = make_float2(0.f, 0.f);
        }
        const void* blk = a.fmt == UGSDR_FMT_CI16 ? (const void*)(static_cast<const short2*>(a.iq) + b * a.spm)
                          : a.fmt == UGSDR_FMT_CI8 ? (const void*)(static_cast<const char2*>(a.iq) + b * a.spm)
                                                   : (const void*)(static_cast<const float2*>(a.iq) + b * a.spm);
        int s = 0;
        for (int64_t n = threadIdx.x; n < a.spm; n += kBankT, ++s) {
            const float2 x = bank_load(blk, a.fmt, n);
            if (s % kReseed == 0) {
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    double ph = fma((double)n, f[c] / a.fs, theta[c]);
                    ph -= rint(ph);
                    float sn, cs;
                    sincospif(2.0f * (float)ph, &sn, &cs);
                    z[c] = make_float2(cs, -sn);
                }
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const float2 m = make_float2(fmaf(x.x, z[c].x, -x.y * z[c].y), fmaf(x.x, z[c].y, x.y * z[c].x));
                z[c] = make_float2(fmaf(z[c].x, w[c].x, -z[c].y * w[c].y), fmaf(z[c].x, w[c].y, z[c].y * w[c].x));
                const uint32_t bi = (uint32_t)(u[c] >> 44);
                const uint32_t bw = min(bi, bi - Lu);
                const uint32_t fr = (uint32_t)(u[c] >> 12);
                const signed char* cc = code + c * LT + bw;
                u[c] += inc[c];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const uint32_t sum = fr + tf[t];
                    const float cv = (float)cc[ti[t] + (sum < fr ? 1 : 0)];
                    acc[c][t].x = fmaf(m.x, cv, acc[c][t].x);
                    acc[c][t].y = fmaf(m.y, cv, acc[c][t].y);
                }
            }
        }
        // workgroup sums of every (channel, tap)
        const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const float sx = wave_sum(acc[c][t].x), sy = wave_sum(acc[c][t].y);
                if (l == 0) red[(wv * CH + c) * NT + t] = make_float2(sx, sy);
            }
        __syncthreads();
        if (threadIdx.x < CH * NT) {
            const int c = threadIdx.x / NT, t = threadIdx.x - c * NT;
            if (c < nch && t < a.n_taps) {
                float2 sum = make_float2(0.f, 0.f);
                for (int i = 0; i < kBankT / 64; ++i) sum = cadd(sum, red[(i * CH + c) * NT + t]);
                a.out[((int64_t)b * a.n_ch + c0 + c) * a.n_taps + t] = sum;
            }
        }
        __syncthreads();
    }
}

// NCO mode for taps a quarter chip apart (the E / P / L and VE / VL taps,
// listed in either direction): with the code expanded to one entry per
// quarter chip, floor(u + k/4) = floor((floor(4u) + k) / 4) for integer k,
// so tap t reads the entry at the sample's quarter index plus a
// compile-time offset: no per-tap index arithmetic.  One channel per
// workgroup of kQSub sub-groups of kBankT threads sharing the channel's
// table; each sub-group correlates its own block at a time, kBankBB blocks
// each.  A thread takes the block's samples n = tid + s T (slot s < 256).
//
// Table: int8 entries, entry i holding quarter index (i - M) mod 4L (M, a
// multiple of 4, pads the start for the taps below the sample's index), and
// a.q_entries - M - 4L entries past the period that let the index run up to
// kQWrap slots beyond it between wrap checks.  Each aligned 32-bit word
// holds one chip's four entries, so the table is written a chip per store.
//
// Code NCO: a 32.32 fixed-point entry index per thread (the integer half is
// the table position of the thread's lowest tap, so the taps' reads are
// immediates off it), seeded per block from the FP64 code phase and
// advanced by round(T 4 dr 2^32) per slot: one 64-bit add per sample, drift
// below 2^-32 entry per slot.  Every kQWrap slots an index past the period
// steps back by 4L.
//
// Carrier: a thread's samples are mixed with W[s] = e^(-j 2 pi T df s), a
// per-channel table built in FP64 (one LDS read per slot, the same address
// across the wave), and the block's tap sums are rotated once by
// z0 = e^(-j 2 pi phase(b spm + tid)) = E_b e^(-j 2 pi tid df):
//   sum_s c x_s e^(-j 2 pi phase(n_s)) = z0 sum_s c x_s W[s].
// The per-block seeds (E_b and the block's code index) come from FP64
// phases computed once per (channel, block) by bank_quarter_seed.
struct QSeed {
    float2 e;     // e^(-j 2 pi phase) of the block's first sample
    uint64_t u;   // 32.32 table position of its lowest tap
};

// constant-memory view of the per-block seeds (read through the scalar
// cache, the same entry across a sub-group) and LDS view of the table, with
// an LDS pointer's 32-bit address (device pass only)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const QSeed cqseed;
typedef __attribute__((address_space(3))) const signed char lds_i8;
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)p;
}
#else
typedef const QSeed cqseed;
typedef const signed char lds_i8;
__host__ __device__ inline uint32_t lds_addr(const void*) { return 0; }
#endif

// the lowest tap's offset in quarter chips (taps listed in direction D)
__host__ __device__ inline int quarter_lo(const BankArgs& a, int nt, int d) {
    return (int)llrint(4.0 * a.taps[d > 0 ? 0 : nt - 1]);
}

__global__ __launch_bounds__(256) void bank_quarter_seed(BankArgs a, int nt, int d, QSeed* seeds) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n_ch * a.n_blocks) return;
    const int c = i / a.n_blocks, b = i - c * a.n_blocks;
    const double f = a.ch[4 * c], th0 = a.ch[4 * c + 1], rate = a.ch[4 * c + 2], tau0 = a.ch[4 * c + 3];
    const double L = (double)a.code_len, tb = (double)b * ((double)a.spm / a.fs);
    double theta = fma(f, tb, th0);
    theta -= rint(theta);
    double sn, co;
    sincospi(2.0 * theta, &sn, &co);
    double tau = fmod(fma(rate, tb, tau0), L);
    tau += tau < 0 ? L : 0.0;
    const double q = 4.0 * tau + (double)(a.tap_margin + quarter_lo(a, nt, d));
    seeds[i] = QSeed{make_float2((float)co, (float)-sn), (uint64_t)llrint(q * 4294967296.0)};
}

template <class S>
__device__ __forceinline__ float2 load_iq(const S* p, int64_t i) {
    const S v = p[i];
    return make_float2((float)v.x, (float)v.y);
}

template <int NT, int D, class S>
__global__ __launch_bounds__(kQSub * kBankT) void bank_nco_quarter(BankArgs a) {
    constexpr int TQ = kQSub * kBankT;
    extern __shared__ __align__(16) unsigned char smem[];
    const int Li = (int)a.code_len, M = a.tap_margin;
    const int LQ = 4 * Li, LT = a.q_entries;
    const int c = blockIdx.x;
    const int sub = threadIdx.x / kBankT, tid = threadIdx.x - sub * kBankT;
    signed char* cq = reinterpret_cast<signed char*>(smem);
    float2* red = reinterpret_cast<float2*>(smem + LT);
    float2* W = red + kQSub * (kBankT / 64) * NT;
    {
        // word w: the four entries of chip (w - M/4) mod L
        const signed char* code = a.codes + (int64_t)c * a.code_len;
        const int step = TQ % Li;
        int chip = ((int)threadIdx.x - M / 4) % Li;
        chip += chip < 0 ? Li : 0;
        for (int w = threadIdx.x; w < LT / 4; w += TQ) {
            reinterpret_cast<uint32_t*>(cq)[w] = (uint32_t)(uint8_t)code[chip] * 0x01010101u;
            chip += step;
            chip -= chip >= Li ? Li : 0;
        }
    }
    const double df = a.ch[4 * c] / a.fs, dr = a.ch[4 * c + 2] / a.fs;
    const int n_w = (int)((a.spm + kBankT - 1) / kBankT);
    for (int s = threadIdx.x; s < n_w; s += TQ) {
        double st = (double)kBankT * df * (double)s;
        st -= rint(st);
        double sn, co;
        sincospi(2.0 * st, &sn, &co);
        W[s] = make_float2((float)co, (float)-sn);
    }
    // the thread's offsets from the block's first sample: carrier phasor and
    // code index
    float2 zt;
    {
        double ph = (double)tid * df;
        ph -= rint(ph);
        double sn, co;
        sincospi(2.0 * ph, &sn, &co);
        zt = make_float2((float)co, (float)-sn);
    }
    // (the integer half also carries the table's LDS address, so a sample's
    // tap reads take it as their address as it is)
    const uint64_t ut = (uint64_t)llrint(4.0 * (double)tid * dr * 4294967296.0) +
                        ((uint64_t)lds_addr(cq) << 32);
    const uint64_t inc = (uint64_t)llrint((double)kBankT * 4.0 * dr * 4294967296.0);
    const uint32_t wrap_at = lds_addr(cq) + (uint32_t)(M + quarter_lo(a, NT, D) + LQ);
    const uint64_t wrap_step = (uint64_t)LQ << 32;
    const int n_slots = a.spm > tid ? (int)((a.spm - tid + kBankT - 1) / kBankT) : 0;
    const cqseed* seeds = (const cqseed*)(a.q_seed + (int64_t)c * a.n_blocks);
    const int wv = tid >> 6, l = tid & 63;
    __syncthreads();
    const int b0 = blockIdx.y * (kQSub * kBankBB);
    for (int i = 0; i < kBankBB; ++i) {
        if (b0 + i * kQSub >= a.n_blocks) break;   // the same for every sub-group
        const int b = b0 + i * kQSub + sub;
        float2 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = make_float2(0.f, 0.f);
        if (b < a.n_blocks) {
            const QSeed sd = seeds[b];
            const float2 z0 = cmul(sd.e, zt);
            uint64_t u = sd.u + ut;
            float2 g[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) g[t] = make_float2(0.f, 0.f);
            const S* xp = static_cast<const S*>(a.iq) + (int64_t)b * a.spm + tid;
            auto step = [&](float2 x, int s) {
                const float2 ws = W[s];
                const float2 m = make_float2(fmaf(x.x, ws.x, -x.y * ws.y), fmaf(x.x, ws.y, x.y * ws.x));
                const lds_i8* p = (const lds_i8*)(uintptr_t)(uint32_t)(u >> 32);
                u += inc;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const float cv = (float)p[D > 0 ? t : NT - 1 - t];
                    g[t].x = fmaf(m.x, cv, g[t].x);
                    g[t].y = fmaf(m.y, cv, g[t].y);
                }
            };
            for (int s = 0; s < n_slots; s += kQWrap) {
                if ((uint32_t)(u >> 32) >= wrap_at) u -= wrap_step;
                if (s + kQWrap <= n_slots) {
                    // four samples' loads in flight before their arithmetic
#pragma unroll
                    for (int h = 0; h < kQWrap; h += 4) {
                        float2 x[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) x[v] = load_iq(xp, (h + v) * kBankT);
#pragma unroll
                        for (int v = 0; v < 4; ++v) step(x[v], s + h + v);
                    }
                    xp += kQWrap * kBankT;
                } else {
                    for (int r = s; r < n_slots; ++r, xp += kBankT) step(load_iq(xp, 0), r);
                }
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = cmul(z0, g[t]);
        }
        float2* rs = red + sub * (kBankT / 64) * NT;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float sx = wave_sum(acc[t].x), sy = wave_sum(acc[t].y);
            if (l == 0) rs[wv * NT + t] = make_float2(sx, sy);
        }
        __syncthreads();
        if (tid < NT && b < a.n_blocks) {
            float2 sum = make_float2(0.f, 0.f);
            for (int k = 0; k < kBankT / 64; ++k) sum = cadd(sum, rs[k * NT + tid]);
            a.out[((int64_t)b * a.n_ch + c) * NT + tid] = sum;
        }
        __syncthreads();
    }
}

template <int NT, int D>
static void launch_quarter(const BankArgs& a, dim3 g, size_t lds, hipStream_t st) {
    const dim3 T(kQSub * kBankT);
    if (a.fmt == UGSDR_FMT_CI16) hipLaunchKernelGGL((bank_nco_quarter<NT, D, short2>), g, T, lds, st, a);
    else if (a.fmt == UGSDR_FMT_CI8) hipLaunchKernelGGL((bank_nco_quarter<NT, D, char2>), g, T, lds, st, a);
    else hipLaunchKernelGGL((bank_nco_quarter<NT, D, float2>), g, T, lds, st, a);
}

template <int NT, int D>
static hipError_t quarter_lds(size_t lds) {
    for (const void* k : {reinterpret_cast<const void*>(bank_nco_quarter<NT, D, short2>),
                          reinterpret_cast<const void*>(bank_nco_quarter<NT, D, char2>),
                          reinterpret_cast<const void*>(bank_nco_quarter<NT, D, float2>)})
        if (hipError_t e = set_lds_limit(k, lds)) return e;
    return hipSuccess;
}

__global__ __launch_bounds__(kBankT) void bank_split_kernel(BankArgs a) {
    const int c = blockIdx.x, b = blockIdx.y;  // the channels of one block run together (IQ block reused from L2)
    __shared__ float2 red[(kBankT / 64) * 2 * kMaxTaps];
    __shared__ float2 part[2 * kMaxTaps];
    const double fc = a.carrier[2 * c];
    // phase after b UpdatePhase() calls (tracking_parameters.hpp:163-165)
    const double ph_b = a.carrier[2 * c + 1] + (double)b * (2.0 * kPi * fc / 1000.0);
    // in-place Translate(x, fs, -fc, -ph_b): FixPhase, no-op for f == 0, fold f < 0
    double phase = fmod(-ph_b, 2.0 * kPi);
    if (phase < 0) phase += 2.0 * kPi;
    double f = -fc;
    const bool mix = fabs(f) != 0.0;
    if (f < 0) f = a.fs + f;
    const double df = f / a.fs, ph_turns = phase / (2.0 * kPi);
    const float* code3 = a.code3 + (int64_t)c * a.code3_len;
    const SplitTap* taps = a.split + (int64_t)c * a.n_taps;
    float2 acc[2 * kMaxTaps];
#pragma unroll
    for (int t = 0; t < 2 * kMaxTaps; ++t) acc[t] = make_float2(0.f, 0.f);
    const void* blk = a.fmt == UGSDR_FMT_CI16 ? (const void*)(static_cast<const short2*>(a.iq) + b * a.spm)
                      : a.fmt == UGSDR_FMT_CI8 ? (const void*)(static_cast<const char2*>(a.iq) + b * a.spm)
                                               : (const void*)(static_cast<const float2*>(a.iq) + b * a.spm);
    for (int64_t n = threadIdx.x; n < a.spm; n += kBankT) {
        float2 x = bank_load(blk, a.fmt, n);
        if (mix) {
            double t = fma((double)n, df, ph_turns);
            t -= rint(t);
            float s, co;
            sincosf((float)(2.0 * kPi * t), &s, &co);
            x = cmul(x, make_float2(co, s));
        }
#pragma unroll
        for (int t = 0; t < kMaxTaps; ++t) {
            if (t >= a.n_taps) break;
            const SplitTap tp = taps[t];
            const bool first = n < tp.first_len;
            const int64_t ci = first ? tp.first_phase + n : tp.second_phase + (n - tp.first_len);
            const float cv = code3[ci];
            float2& ac = acc[2 * t + (first ? 0 : 1)];
            ac.x = fmaf(x.x, cv, ac.x);
            ac.y = fmaf(x.y, cv, ac.y);
        }
    }
    bank_reduce_store<2 * kMaxTaps>(acc, 2 * a.n_taps, part, red);
    __syncthreads();
    if (threadIdx.x < a.n_taps) {
        // AddWithPhase (tracking_parameters.hpp:89-98)
        const SplitTap tp = taps[threadIdx.x];
        float2 lhs = part[2 * threadIdx.x], rhs = part[2 * threadIdx.x + 1];
        if (tp.rel < 0.5) {
            const float2 tmp = lhs;
            lhs = rhs;
            rhs = tmp;
        }
        const double wl = (double)lhs.x / tp.rel, wr = (double)rhs.x / (1.0 - tp.rel);
        const float2 r = fabs(wl + wr) > fabs(wl) ? cadd(lhs, rhs) : csub(lhs, rhs);
        a.out[((int64_t)b * a.n_ch + c) * a.n_taps + threadIdx.x] = r;
    }
}

// CorrelateSplit index arithmetic (tracking_parameters.hpp:475-486)
static SplitTap split_tap(double phase, int64_t spm, double code_period_ms) {
    const double cps = (double)spm * code_period_ms;
    while (phase < 0) phase += cps;
    if (phase > cps) phase = std::fmod(phase, cps);
    SplitTap t{};
    t.first_len = (int)((size_t)std::ceil(phase) % (size_t)spm);
    t.first_phase = (int)(size_t)std::ceil(2 * cps - phase);
    t.second_phase = t.first_phase + t.first_len;
    t.rel = std::fmod(phase, (double)spm) / (double)spm;
    return t;
}

int bank_impl(ugsdr_hip_ctx* ctx, const ugsdr_bank_request* r, ugsdr_channel_state* ch, float* out) {
    UGSDR_CHECK_ARG(r && ch && out, "null argument");
    UGSDR_CHECK_ARG(r->iq && r->spm > 0 && r->n_blocks > 0 && r->fs > 0, "bad signal");
    if (r->n_blocks > 65535) {
        set_error("at most 65535 blocks per call");
        return UGSDR_ERR_SIZE;
    }
    UGSDR_CHECK_ARG(r->n_channels > 0 && r->n_taps > 0 && r->n_taps <= kMaxTaps && r->taps_chips, "bad taps");
    UGSDR_CHECK_ARG(r->codes && r->code_len > 0, "bad codes");
    UGSDR_CHECK_ARG(r->fmt == UGSDR_FMT_CI8 || r->fmt == UGSDR_FMT_CI16 || r->fmt == UGSDR_FMT_CF32, "bad format");
    UGSDR_CHECK_ARG(r->mode == UGSDR_BANK_NCO || r->mode == UGSDR_BANK_REF_SPLIT, "bad mode");
    const int64_t nsamp = r->spm * r->n_blocks;
    const size_t bps = r->fmt == UGSDR_FMT_CI8 ? 2 : r->fmt == UGSDR_FMT_CI16 ? 4 : 8;
    const void* iq = r->iq;
    if (!r->iq_on_device) {
        UGSDR_HIP_TRY(ctx->iq.reserve(nsamp * bps));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->iq.ptr, r->iq, nsamp * bps, hipMemcpyHostToDevice, ctx->stream));
        iq = ctx->iq.ptr;
    }
    BankArgs a{};
    a.iq = iq;
    a.fmt = r->fmt;
    a.spm = r->spm;
    a.n_blocks = (int)r->n_blocks;
    a.n_ch = r->n_channels;
    a.n_taps = r->n_taps;
    a.fs = r->fs;
    for (int t = 0; t < r->n_taps; ++t) a.taps[t] = r->taps_chips[t];
    const size_t out_bytes = (size_t)r->n_blocks * r->n_channels * r->n_taps * sizeof(float2);
    UGSDR_HIP_TRY(ctx->results.reserve(out_bytes));
    a.out = ctx->results.as<float2>();
    const dim3 grid((unsigned)r->n_channels, (unsigned)r->n_blocks);
    if (r->mode == UGSDR_BANK_NCO) {
        std::vector<double> chs(4 * (size_t)r->n_channels);
        for (int c = 0; c < r->n_channels; ++c) {
            chs[4 * c] = ch[c].carrier_frequency;
            chs[4 * c + 1] = ch[c].carrier_phase;
            chs[4 * c + 2] = ch[c].code_rate;
            chs[4 * c + 3] = ch[c].code_phase;
        }
        const size_t code_bytes = (size_t)r->n_channels * r->code_len;
        // the quarter-table kernel's per-block seeds follow the codes
        const size_t seed_off = (chs.size() * 8 + code_bytes + 15) / 16 * 16;
        UGSDR_HIP_TRY(ctx->tables.reserve(seed_off + (size_t)r->n_channels * r->n_blocks * sizeof(QSeed)));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->tables.ptr, chs.data(), chs.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->tables.as<char>() + chs.size() * 8, r->codes, code_bytes,
                                     hipMemcpyHostToDevice, ctx->stream));
        a.ch = ctx->tables.as<double>();
        a.codes = reinterpret_cast<const signed char*>(ctx->tables.as<char>() + chs.size() * 8);
        a.code_len = r->code_len;
        // wrap padding: the largest |floor(tap)| + 1 chips on each side
        double reach = 0;
        for (int t = 0; t < r->n_taps; ++t) reach = std::max(reach, std::fabs(std::floor(r->taps_chips[t])) + 1);
        UGSDR_CHECK_ARG(reach < (double)r->code_len, "taps reach beyond one code period");
        // The multi-channel kernel keeps a block's code phase within one
        // period plus the block's own span, so it needs positive code rates,
        // at most one code period (plus 2 chips) per block and a 20-bit chip
        // count; other channel sets run the one-channel kernel.
        double span_max = 0;
        bool multi = r->code_len < (1 << 19);
        for (int c = 0; c < r->n_channels; ++c) {
            multi = multi && ch[c].code_rate > 0;
            span_max = std::max(span_max, (double)r->spm * ch[c].code_rate / r->fs);
        }
        multi = multi && span_max <= (double)r->code_len + 2.0;
        const int margin_multi = (int)reach + 3;
        const int nt = r->n_taps <= 3 ? 3 : r->n_taps <= 5 ? 5 : kMaxTaps;
        const size_t lds_multi = ((size_t)kBankCH * (r->code_len + 2 * margin_multi) + 15) / 16 * 16 +
                                 (size_t)(kBankT / 64) * kBankCH * nt * sizeof(float2);
        multi = multi && lds_multi <= 64 * 1024;
        // taps a quarter chip apart (4 o_t = 4 o_0 + t D, D = +-1; three or
        // five of them) take the quarter-table kernel where its table (about
        // 4 L bytes) leaves room for three workgroups per CU and a block is
        // at most kQSlots slots
        int dir = 0;
        if (r->n_taps == 3 || r->n_taps == 5) {
            const double d = 4.0 * (r->taps_chips[1] - r->taps_chips[0]);
            dir = std::fabs(d - 1.0) < 1e-9 ? 1 : std::fabs(d + 1.0) < 1e-9 ? -1 : 0;
            const double q0 = 4.0 * r->taps_chips[0];
            if (std::fabs(q0 - std::nearbyint(q0)) > 1e-9) dir = 0;
            for (int t = 0; t < r->n_taps && dir; ++t)
                if (std::fabs(4.0 * r->taps_chips[t] - (q0 + t * dir)) > 1e-9) dir = 0;
        }
        double tap_abs = 0;
        for (int t = 0; t < r->n_taps; ++t) tap_abs = std::max(tap_abs, std::fabs(r->taps_chips[t]));
        const int margin_q = 4 * (int)std::ceil(tap_abs) + 8;
        // quarter table: the period, the start margin, and past the period
        // the margin plus kQWrap slots of index advance (the fastest
        // channel's whole entries per slot, plus its carry) and the taps
        double dr_max = 0;
        for (int c = 0; c < r->n_channels; ++c) dr_max = std::max(dr_max, ch[c].code_rate / r->fs);
        const int64_t q_entries =
            (4 * r->code_len + 2 * margin_q + kQWrap * ((int64_t)std::floor(kBankT * 4.0 * dr_max) + 1) + r->n_taps +
             15) / 16 * 16;
        const int64_t n_w = (r->spm + kBankT - 1) / kBankT;
        const size_t lds_q =
            (size_t)q_entries + (size_t)(kQSub * (kBankT / 64) * r->n_taps + n_w) * sizeof(float2);
        const bool quarter = multi && dir != 0 && n_w <= kQSlots && lds_q <= (kQSub > 1 ? 156 : 52) * 1024;
        ctx->mark(0);
        if (quarter) {
            a.tap_margin = margin_q;
            a.q_entries = (int)q_entries;
            QSeed* seeds = reinterpret_cast<QSeed*>(ctx->tables.as<char>() + seed_off);
            const int n_seed = r->n_channels * (int)r->n_blocks;
            hipLaunchKernelGGL(bank_quarter_seed, dim3((unsigned)((n_seed + 255) / 256)), dim3(256), 0, ctx->stream, a,
                               (int)r->n_taps, dir, seeds);
            a.q_seed = seeds;
            const dim3 g((unsigned)r->n_channels,
                         (unsigned)((r->n_blocks + kQSub * kBankBB - 1) / (kQSub * kBankBB)));
            if (r->n_taps == 3) {
                UGSDR_HIP_TRY((dir > 0 ? quarter_lds<3, 1>(lds_q) : quarter_lds<3, -1>(lds_q)));
                if (dir > 0) launch_quarter<3, 1>(a, g, lds_q, ctx->stream);
                else launch_quarter<3, -1>(a, g, lds_q, ctx->stream);
            } else {
                UGSDR_HIP_TRY((dir > 0 ? quarter_lds<5, 1>(lds_q) : quarter_lds<5, -1>(lds_q)));
                if (dir > 0) launch_quarter<5, 1>(a, g, lds_q, ctx->stream);
                else launch_quarter<5, -1>(a, g, lds_q, ctx->stream);
            }
        } else if (multi) {
            a.tap_margin = margin_multi;
            const dim3 g((unsigned)((r->n_channels + kBankCH - 1) / kBankCH),
                         (unsigned)((r->n_blocks + kBankBB - 1) / kBankBB));
            const void* k = nt == 3 ? reinterpret_cast<const void*>(bank_nco_multi<3>)
                            : nt == 5 ? reinterpret_cast<const void*>(bank_nco_multi<5>)
                                      : reinterpret_cast<const void*>(bank_nco_multi<kMaxTaps>);
            UGSDR_HIP_TRY(set_lds_limit(k, lds_multi));
            if (nt == 3) hipLaunchKernelGGL(bank_nco_multi<3>, g, dim3(kBankT), lds_multi, ctx->stream, a);
            else if (nt == 5) hipLaunchKernelGGL(bank_nco_multi<5>, g, dim3(kBankT), lds_multi, ctx->stream, a);
            else hipLaunchKernelGGL(bank_nco_multi<kMaxTaps>, g, dim3(kBankT), lds_multi, ctx->stream, a);
        } else {
            a.tap_margin = (int)reach;
            const size_t lds =
                ((r->code_len + 2 * a.tap_margin + 15) / 16) * 16 + (kBankT / 64) * kMaxTaps * sizeof(float2);
            if (lds > 160 * 1024) {
                set_error("code too long for the LDS code table");
                return UGSDR_ERR_UNSUPPORTED;
            }
            UGSDR_HIP_TRY(set_lds_limit(reinterpret_cast<const void*>(bank_nco_kernel), lds));
            hipLaunchKernelGGL(bank_nco_kernel, grid, dim3(kBankT), lds, ctx->stream, a);
        }
        UGSDR_HIP_TRY(hipGetLastError());
        ctx->mark(1);
    } else {
        UGSDR_CHECK_ARG(r->samples_per_chip > 0 && r->code_period_ms > 0, "bad replica rate");
        // 3-period replica table, Codes::FillCodesImpl (src/tracking/tracker.hpp:39-51)
        const int64_t cps = (int64_t)std::llround((double)r->spm * r->code_period_ms);
        const int64_t len3 = 3 * cps;
        const int64_t chips3 = 3 * r->code_len;
        std::vector<float> code3((size_t)r->n_channels * len3);
        for (int c = 0; c < r->n_channels; ++c)
            for (int64_t i = 0; i < len3; ++i)
                code3[(size_t)c * len3 + i] =
                    (float)r->codes[(size_t)c * r->code_len + ((i * chips3 / len3) % chips3) % r->code_len];
        std::vector<SplitTap> split((size_t)r->n_channels * r->n_taps);
        std::vector<double> carr(2 * (size_t)r->n_channels);
        for (int c = 0; c < r->n_channels; ++c) {
            carr[2 * c] = ch[c].carrier_frequency;
            carr[2 * c + 1] = ch[c].carrier_phase;
            for (int t = 0; t < r->n_taps; ++t)
                split[(size_t)c * r->n_taps + t] =
                    split_tap(ch[c].code_phase + r->samples_per_chip * r->taps_chips[t], r->spm, r->code_period_ms);
        }
        const size_t b1 = code3.size() * 4, b2 = split.size() * sizeof(SplitTap), b3 = carr.size() * 8;
        UGSDR_HIP_TRY(ctx->tables.reserve(b1 + b2 + b3 + 64));
        char* d = ctx->tables.as<char>();
        UGSDR_HIP_TRY(hipMemcpyAsync(d, code3.data(), b1, hipMemcpyHostToDevice, ctx->stream));
        const size_t o2 = (b1 + 15) / 16 * 16, o3 = o2 + (b2 + 15) / 16 * 16;
        UGSDR_HIP_TRY(hipMemcpyAsync(d + o2, split.data(), b2, hipMemcpyHostToDevice, ctx->stream));
        UGSDR_HIP_TRY(hipMemcpyAsync(d + o3, carr.data(), b3, hipMemcpyHostToDevice, ctx->stream));
        a.code3 = reinterpret_cast<const float*>(d);
        a.code3_len = len3;
        a.split = reinterpret_cast<const SplitTap*>(d + o2);
        a.carrier = reinterpret_cast<const double*>(d + o3);
        ctx->mark(0);
        hipLaunchKernelGGL(bank_split_kernel, grid, dim3(kBankT), 0, ctx->stream, a);
        UGSDR_HIP_TRY(hipGetLastError());
        ctx->mark(1);
    }
    UGSDR_HIP_TRY(hipMemcpyAsync(out, a.out, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->profile) {
        // the bank kernel's time goes in the correlation slot
        ctx->last_ms[0] = ctx->last_ms[1] = ctx->last_ms[3] = 0.f;
        UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[2], ctx->ev[0], ctx->ev[1]));
    }
    // advance the channel states past the processed blocks
    for (int c = 0; c < r->n_channels; ++c) {
        if (r->mode == UGSDR_BANK_NCO) {
            const double span = (double)nsamp / r->fs;
            double th = ch[c].carrier_phase + ch[c].carrier_frequency * span;
            ch[c].carrier_phase = th - std::floor(th);
            double tau = ch[c].code_phase + ch[c].code_rate * span;
            tau = std::fmod(tau, (double)r->code_len);
            if (tau < 0) tau += (double)r->code_len;
            ch[c].code_phase = tau;
        } else {
            for (int64_t b = 0; b < r->n_blocks; ++b)
                ch[c].carrier_phase += 2 * M_PI * ch[c].carrier_frequency / 1000.0;
        }
    }
    return UGSDR_OK;
}

}  // namespace ugsdr_hip

extern "C" int ugsdr_hip_correlate_bank(ugsdr_hip_ctx* ctx, const ugsdr_bank_request* req, ugsdr_channel_state* ch,
                                        float* out) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return ugsdr_hip::bank_impl(ctx, req, ch, out);
}
