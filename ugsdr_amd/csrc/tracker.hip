// GPU-resident tracking loops (Tracker::Track, src/tracking/tracker.hpp:156-197).
//
// One workgroup per channel walks the epochs of a call in order.  Per 1 ms
// epoch:
//   - wave 0 has planned it: the code phase of the epoch (TrackSingleSatellite,
//     tracker.hpp:165), the early / prompt / late split points and replica
//     offsets (CorrelateSplit, tracking_parameters.hpp:471-486) and the tone
//     of the in-place Mixer::Translate (mixer.hpp:33-41), after which the
//     carrier phase advances (UpdatePhase, tracking_parameters.hpp:163-165);
//   - every thread mixes groups of four samples and accumulates the three
//     taps against the 3-period replica table, a snapshot of each running sum
//     marking the tap's split point;
//   - the sums are reduced over each wave and, in LDS, over the waves; wave 0
//     joins each tap's two spans (AddWithPhase, tracking_parameters.hpp:89-98),
//     runs the PLL/FLL and DLL updates (:495-528) in FP64 and plans the next
//     epoch, so no epoch waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "acq.hpp"
#include "ctx.hpp"
#include "dev_fft.hpp"
#include "formats.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

// workgroup sizes: 1024 threads per channel while every channel has a CU of
// its own, 512 (two channels per CU) up to twice as many channels as CUs,
// 256 (four per CU) beyond; each CU then overlaps one channel's serial loop
// step with the others' correlations
constexpr int kTrackT = 1024;

// Loop filter gains of TrackingParameters (tracking_parameters.hpp:72-87):
// PLL noise bandwidth 25 Hz aided by a 250 Hz FLL, DLL noise bandwidth 0.7 Hz,
// damping 1, gain 0.1, 1 ms summation.
namespace gains {
constexpr double kPllWn = 25.0 / 0.53;
constexpr double kPll1 = 0.001 * (kPllWn * kPllWn) + 1.414 * kPllWn;
constexpr double kPll2 = 1.414 * kPllWn;
constexpr double kPll3 = 0.001 * (250.0 / 0.25);
constexpr double kDllWn = 0.7 * 8.0 * 1.0 / (4.0 * 1.0 * 1.0 + 1.0);
constexpr double kDllTau1 = 0.1 / (kDllWn * kDllWn);
constexpr double kDllTau2 = 2.0 * 1.0 / kDllWn;
constexpr double kDll1 = kDllTau2 / kDllTau1;
constexpr double kDll2 = 0.001 / kDllTau1;
}  // namespace gains

struct TrackArgs {
    SampleSource src;
    int64_t spm;
    int n_blocks;
    int n_ch;
    double fs;
    double spacing;              // early/late offset in chips
    const float* tabs;           // per channel: 4 alignment copies of the 3-period replica
    int64_t len3p;               // floats per copy (multiple of 4)
    ugsdr_track_state* state;    // n_ch
    float2* epl;                 // n_blocks x n_ch x 3
    double* trace;               // n_blocks x n_ch x 4, or null
};

// What every thread needs to correlate one epoch, written by wave 0.
struct EpochPlan {
    int fl[3];        // first-span length of each tap (CorrelateSplit)
    int fa[3];        // 4-aligned table offset of the tap: fp - (fp & 3), fp = ceil(2P - phase)
    int al[3];        // fp & 3: the alignment copy that makes fa + 4g a 16-byte boundary
    double rel[3];    // fractional position of the split in the epoch (AddWithPhase)
    double df, pt;    // mixing tone: turns per sample and start phase in turns (df = 0: no mixing)
    float2 w1, wg;    // e^{j 2 pi df}, e^{j 2 pi 4 T df}
};

// Constants of a channel's loop, set once per launch (the reciprocals turn
// the per-epoch divisions by constants into multiplications)
struct ChanConst {
    double fs_ms;       // fs / 1e3
    double half_fs_ms;  // fs / 1000 / 2
    double off;         // GetSamplesPerChip() * spacing
    double base_ratio;  // base code frequency / 1.023e6
    double inv_base;    // 1 / base code frequency
    double inv_fs;      // 1 / fs
    double fs;
    int period;         // code period in ms
};

constexpr double kInv2Pi = 1.0 / (2.0 * kPi), kInvPi = 1.0 / kPi;

__device__ __forceinline__ double lane_value(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// TrackingParameters::Pll (tracking_parameters.hpp:495-511).  The prompt is
// complex<float>: the I/Q ratio and the cross / dot products are float, the
// discriminators and the state FP64.
__device__ __forceinline__ void pll_update(ugsdr_track_state& s, float2 p) {
    const double e = p.x != 0.f ? atan((double)(p.y / p.x)) * kInv2Pi : 0.0;
    const float qx = s.previous_prompt[0], qy = s.previous_prompt[1];
    const float cross = p.x * qy - qx * p.y;
    const float dot = fabsf(p.x * qx + qy * p.y);
    const double ferr = (double)atan2f(cross, dot) * kInvPi;
    s.carrier_frequency += gains::kPll1 * e - gains::kPll2 * s.carrier_phase_error - gains::kPll3 * ferr;
    s.carrier_phase_error = e;
    s.carrier_phase += e * kPi / 2;
    s.previous_prompt[0] = p.x;
    s.previous_prompt[1] = p.y;
}

// TrackingParameters::Dll (tracking_parameters.hpp:513-528): normalised
// early-minus-late envelope in float, NCO and code phase in FP64
__device__ __forceinline__ void dll_update(ugsdr_track_state& s, float2 e, float2 l, const ChanConst& cc) {
    const float ea = hypotf(e.x, e.y), la = hypotf(l.x, l.y);
    double err = 0.0;
    if (ea + la != 0.f) err = (ea - la) / (ea + la);
    s.code_nco += (gains::kDll1 * (err - s.code_error) + gains::kDll2 * err) * cc.base_ratio;
    s.code_error = err;
    s.code_frequency = s.base_code_frequency - s.code_nco;
    s.code_phase -= cc.half_fs_ms * (s.code_frequency * cc.inv_base - 1);
}

// Wave 0 between epochs, all 64 lanes.  Unless `first`: lane t joins tap
// t's two spans over the waves (AddWithPhase, tracking_parameters.hpp:89-98)
// and lane 0 runs the PLL/FLL and DLL updates and the trace.  Unless
// `last`: the next epoch is planned - TrackSingleSatellite's code phase
// (tracker.hpp:165) and the tone of Translate(x, fs, -f, -phase)
// (mixer.hpp:33-41) on lane 0, then UpdatePhase (tracking_parameters.hpp:
// 163-165); CorrelateSplit's split of tap t = early / prompt / late
// (tracking_parameters.hpp:471-486) on lane t; the tone's per-sample and
// per-group phasors on lanes 3 and 4.  The state goes through registers of
// lane 0 and back to LDS once.
template <int T>
__device__ __noinline__ void epoch_tail(ugsdr_track_state& s_lds, EpochPlan& plan, const ChanConst& cc_lds,
                                        const float4* red, float2* o, double* tr, bool first, bool last, int spm,
                                        int lane) {
    const ChanConst cc = cc_lds;
    float2 e = make_float2(0.f, 0.f);
    if (!first && lane < 3) {
        float2 lhs = make_float2(0.f, 0.f), rhs = lhs;
#pragma unroll
        for (int w = 0; w < T / 64; ++w) {
            const float4 v = red[w * 3 + lane];
            lhs = cadd(lhs, make_float2(v.x, v.y));
            rhs = cadd(rhs, make_float2(v.z, v.w));
        }
        const double rel = plan.rel[lane];
        if (rel < 0.5) {
            const float2 tmp = lhs;
            lhs = rhs;
            rhs = tmp;
        }
        const double wl = (double)lhs.x / rel, wr = (double)rhs.x / (1.0 - rel);
        e = fabs(wl + wr) > fabs(wl) ? cadd(lhs, rhs) : csub(lhs, rhs);
        o[lane] = e;
    }
    const float2 early = make_float2(lane_value(e.x, 0), lane_value(e.y, 0));
    const float2 prompt = make_float2(lane_value(e.x, 1), lane_value(e.y, 1));
    const float2 late = make_float2(lane_value(e.x, 2), lane_value(e.y, 2));
    double cp = 0.0, df = 0.0;
    if (lane == 0) {
        ugsdr_track_state s = s_lds;
        if (!first) {
            pll_update(s, prompt);
            dll_update(s, early, late, cc);
            ++s.epochs;
            if (tr) {
                tr[0] = -s.carrier_phase * kInv2Pi;
                tr[1] = s.carrier_frequency;
                tr[2] = s.code_phase;
                tr[3] = s.code_frequency;
            }
        }
        if (!last) {
            const int64_t k = cc.period == 1 ? 0 : s.epochs % cc.period;
            cp = s.code_phase - cc.fs_ms * (double)k;
            double f = -s.carrier_frequency;
            double pt = 0.0;  // f == 0: Translate leaves the epoch as it is
            if (f != 0.0) {
                // -phase reduced to [0, 2 pi) (floor form: within an ulp of fmod's)
                const double x = -s.carrier_phase;
                double phase = x - 2.0 * kPi * floor(x * kInv2Pi);
                if (phase < 0) phase += 2.0 * kPi;
                if (phase >= 2.0 * kPi) phase -= 2.0 * kPi;
                if (f < 0) f = cc.fs + f;
                df = f * cc.inv_fs;
                pt = phase * kInv2Pi;
            }
            plan.df = df;
            plan.pt = pt;
            s.carrier_phase += 2 * kPi * s.carrier_frequency / 1000.0;  // UpdatePhase
        }
        s_lds = s;
    }
    if (last) return;
    cp = lane_value(cp, 0);
    df = lane_value(df, 0);
    if (lane < 3) {
        const double cps = (double)spm * cc.period;
        double p = cp + (double)(1 - lane) * cc.off;  // cp + off, cp, cp - off
        while (p < 0) p += cps;
        if (p > cps) p = fmod(p, cps);
        const int fp = (int)ceil(2 * cps - p);
        int fl = (int)ceil(p);
        while (fl >= spm) fl -= spm;
        // p mod spm, exactly: q spm is exact and p - q spm is exact for
        // p in [q spm, (q + 1) spm)
        const double q = floor(p / spm);
        double r = p - q * spm;
        if (r < 0) r += spm;
        if (r >= spm) r -= spm;
        plan.fl[lane] = fl;
        plan.al[lane] = fp & 3;
        plan.fa[lane] = fp - (fp & 3);
        plan.rel[lane] = r / spm;
    } else if (lane == 3 || lane == 4) {
        float2 w = make_float2(1.f, 0.f);
        if (df != 0.0) {
            double t = lane == 3 ? df : 4.0 * T * df;
            t -= rint(t);
            float sn, cs;
            sincospif((float)(2.0 * t), &sn, &cs);
            w = make_float2(cs, sn);
        }
        if (lane == 3)
            plan.w1 = w;
        else
            plan.wg = w;
    }
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

// Sums the six complex partials of every lane over the wave by halving:
// after the exchanges over lane bits 5 and 4 (shuffles) each row of 16 lanes
// holds three of the twelve components, summed over its partners, and DPP
// exchanges inside the row (xor 1 and xor 2 as quad permutations, then the
// half-row and row mirrors) complete them.  Every lane of row q ends with
// components 3q..3q+2.
__device__ __forceinline__ void wave_reduce12(float (&v)[12], int lane, float (&out)[3]) {
    float h[6];
    const bool b5 = lane & 32;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float keep = b5 ? v[6 + i] : v[i], give = b5 ? v[i] : v[6 + i];
        h[i] = keep + __shfl_xor(give, 32, 64);
    }
    const bool b4 = lane & 16;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float keep = b4 ? h[3 + i] : h[i], give = b4 ? h[i] : h[3 + i];
        out[i] = keep + __shfl_xor(give, 16, 64);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] += dpp_move<0xB1>(out[i]);   // quad_perm [1,0,3,2]
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] += dpp_move<0x4E>(out[i]);   // quad_perm [2,3,0,1]
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] += dpp_move<0x141>(out[i]);  // row_half_mirror
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] += dpp_move<0x140>(out[i]);  // row_mirror
}

template <int FMT, int T>
__global__ __launch_bounds__(T) void track_kernel(TrackArgs a) {
    const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ EpochPlan plan;
    __shared__ float4 red[(T / 64) * 3];
    __shared__ ugsdr_track_state s;
    const float* tabs = a.tabs + (int64_t)c * 4 * a.len3p;
    const int spm = (int)a.spm, nfull = spm >> 2;
    __shared__ ChanConst cc;
    if (tid == 0) {
        s = a.state[c];
        cc.fs = a.fs;
        cc.fs_ms = a.fs / 1e3;
        cc.half_fs_ms = a.fs / 1000 / 2;
        cc.off = a.fs / s.base_code_frequency * a.spacing;
        cc.base_ratio = s.base_code_frequency / 1.023e6;
        cc.inv_base = 1.0 / s.base_code_frequency;
        cc.inv_fs = 1.0 / a.fs;
        cc.period = s.code_period_ms;
    }
    __syncthreads();
    if (wave == 0) epoch_tail<T>(s, plan, cc, red, nullptr, nullptr, true, false, spm, lane);
    __syncthreads();
    for (int b = 0; b < a.n_blocks; ++b) {
        const int64_t base = (int64_t)b * a.spm;
        int fl[3];
        const float4* tab[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            fl[t] = plan.fl[t];
            tab[t] = reinterpret_cast<const float4*>(tabs + plan.al[t] * a.len3p + plan.fa[t]);
        }
        const float2 wg = plan.wg, w1 = plan.w1;
        const double df = plan.df, pt = plan.pt;
        float2 z[4];
        {
            double ph = fma(4.0 * tid, df, pt);
            ph -= rint(ph);
            float sn, cs;
            sincospif((float)(2.0 * ph), &sn, &cs);
            z[0] = df == 0.0 ? make_float2(1.f, 0.f) : make_float2(cs, sn);
            z[1] = cmul(z[0], w1);
            z[2] = cmul(z[1], w1);
            z[3] = cmul(z[2], w1);
        }
        float2 tot[3], snap[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) tot[t] = snap[t] = make_float2(0.f, 0.f);
        // groups of four whole samples, the next group's loads in flight while
        // the current one is summed (the prefetch index is clamped in range)
        float2 xn[4];
        float4 cn[3];
        auto fetch = [&](int g) {
            g = min(g, nfull - 1);
#pragma unroll
            for (int u = 0; u < 4; ++u) xn[u] = load_sample_fmt<FMT>(a.src, base + 4 * g + u);
#pragma unroll
            for (int t = 0; t < 3; ++t) cn[t] = tab[t][g];
        };
        fetch(tid);
        for (int g = tid; g < nfull; g += T) {
            float2 x[4];
            float4 cv[3];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = xn[u];
#pragma unroll
            for (int t = 0; t < 3; ++t) cv[t] = cn[t];
            fetch(g + T);
            const int n0 = 4 * g;
            float2 xm[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xm[u] = cmul(x[u], z[u]);
                z[u] = cmul(z[u], wg);
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const float cu[4] = {cv[t].x, cv[t].y, cv[t].z, cv[t].w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    tot[t].x = fmaf(xm[u].x, cu[u], tot[t].x);
                    tot[t].y = fmaf(xm[u].y, cu[u], tot[t].y);
                }
                if (n0 + 4 <= fl[t]) {
                    snap[t] = tot[t];
                } else if (n0 < fl[t]) {
                    // the split falls inside this group: the first span ends at fl
                    float2 sp = tot[t];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (n0 + u >= fl[t]) {
                            sp.x = fmaf(-xm[u].x, cu[u], sp.x);
                            sp.y = fmaf(-xm[u].y, cu[u], sp.y);
                        }
                    snap[t] = sp;
                }
            }
        }
        // the epoch's last spm mod 4 samples, one per thread; every group of
        // such a thread lies before them, so its snapshot is its running sum
        // whenever the sample itself is in the first span
        if (tid < spm - 4 * nfull) {
            const int n = 4 * nfull + tid;
            float2 xm = load_sample_fmt<FMT>(a.src, base + n);
            if (df != 0.0) {
                double ph = fma((double)n, df, pt);
                ph -= rint(ph);
                float sn, cs;
                sincospif((float)(2.0 * ph), &sn, &cs);
                xm = cmul(xm, make_float2(cs, sn));
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const float cu = tabs[plan.fa[t] + plan.al[t] + n];
                tot[t].x = fmaf(xm.x, cu, tot[t].x);
                tot[t].y = fmaf(xm.y, cu, tot[t].y);
                if (n < fl[t]) snap[t] = tot[t];
            }
        }
        float v[12];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            v[4 * t] = snap[t].x;
            v[4 * t + 1] = snap[t].y;
            v[4 * t + 2] = tot[t].x - snap[t].x;
            v[4 * t + 3] = tot[t].y - snap[t].y;
        }
        float part[3];
        wave_reduce12(v, lane, part);
        if ((lane & 15) == 0) {
            float* r = reinterpret_cast<float*>(red) + wave * 12 + 3 * (lane >> 4);
            r[0] = part[0];
            r[1] = part[1];
            r[2] = part[2];
        }
        __syncthreads();
        if (wave == 0)
            epoch_tail<T>(s, plan, cc, red, a.epl + ((int64_t)b * a.n_ch + c) * 3,
                       a.trace ? a.trace + ((int64_t)b * a.n_ch + c) * 4 : nullptr, false, b + 1 == a.n_blocks, spm,
                       lane);
        __syncthreads();
    }
    // no plan is made past the last epoch: the state is the one after n_blocks
    if (tid == 0) a.state[c] = s;
}

// The 3-period replica table of Codes::FillCodesImpl (tracker.hpp:39-51):
// entry i of the 3 P spm table is chip floor(i L / (P spm)) mod L, the
// floor-index upsampling of the code repeated three times.  Copy a holds
// entry (j + a) mod 3 P spm at j, so that the reads of a tap whose table
// offset is a mod 4 start on 16-byte boundaries.
__global__ void track_tables_kernel(const int8_t* codes, int64_t code_len, int64_t per_period, int64_t len3,
                                    int64_t len3p, int n_ch, float* tabs) {
    const int64_t total = (int64_t)n_ch * 4 * len3p;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = k % len3p, ca = k / len3p;
        const int64_t al = ca & 3, ch = ca >> 2;
        const int64_t i = (j + al) % len3;
        tabs[k] = (float)codes[ch * code_len + (i * code_len / per_period) % code_len];
    }
}

template <int T>
static int launch_track(int fmt, int n_ch, const TrackArgs& a, hipStream_t st) {
    const dim3 grid((unsigned)n_ch), block(T);
    switch (fmt) {
        case UGSDR_FMT_CI8: hipLaunchKernelGGL((track_kernel<UGSDR_FMT_CI8, T>), grid, block, 0, st, a); break;
        case UGSDR_FMT_CI16: hipLaunchKernelGGL((track_kernel<UGSDR_FMT_CI16, T>), grid, block, 0, st, a); break;
        case UGSDR_FMT_CF32: hipLaunchKernelGGL((track_kernel<UGSDR_FMT_CF32, T>), grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL((track_kernel<-1, T>), grid, block, 0, st, a); break;
    }
    return UGSDR_OK;
}

static int track_impl(ugsdr_hip_ctx* ctx, const ugsdr_track_request* r, ugsdr_track_state* ch, float* epl,
                      double* trace) {
    UGSDR_CHECK_ARG(r && ch && epl, "null argument");
    UGSDR_CHECK_ARG(r->iq && r->spm >= 4 && r->n_blocks > 0 && r->fs > 0, "bad signal");
    UGSDR_CHECK_ARG(r->n_channels > 0 && r->codes && r->code_len > 0, "bad channels");
    UGSDR_CHECK_ARG(r->spacing_chips >= 0 && r->spacing_chips < 1, "bad early/late spacing");
    UGSDR_CHECK_ARG(r->n_blocks < (1 << 30) && r->spm < (1 << 24) && r->spm * r->n_blocks < ((int64_t)1 << 31),
                    "at most 2^31 - 1 samples a call");
    const int period = ch[0].code_period_ms;
    for (int c = 0; c < r->n_channels; ++c) {
        UGSDR_CHECK_ARG(ch[c].code_period_ms == period && period >= 1, "channels of one call share a code period");
        UGSDR_CHECK_ARG(ch[c].base_code_frequency > 0, "bad code frequency");
    }
    UGSDR_CHECK_ARG(3 * r->spm * period + 8 < ((int64_t)1 << 31), "replica table beyond 2^31 samples");
    SampleSource src{};
    if (int rc = format_geometry(r->fmt, r->fs, src.spm, src.epoch_bytes)) return rc;
    src.fmt = r->fmt;
    const int64_t nsamp = r->spm * r->n_blocks;
    const int64_t in_bytes = format_bytes(r->fmt, nsamp, src.spm, src.epoch_bytes);
    hipStream_t st = ctx->stream;
    if (!r->iq_on_device) {
        if (r->fmt == UGSDR_FMT_BBP_DDC)
            if (int rc = verify_bbp_headers(r->iq, in_bytes / src.epoch_bytes, src.epoch_bytes)) return rc;
        UGSDR_HIP_TRY(ctx->iq.reserve(in_bytes));
        UGSDR_HIP_TRY(hipMemcpyAsync(ctx->iq.ptr, r->iq, in_bytes, hipMemcpyHostToDevice, st));
        src.p = ctx->iq.ptr;
    } else {
        src.p = r->iq;
    }
    // replica tables, rebuilt on the device when the codes or the rate change
    const int64_t cps = r->spm * period, len3 = 3 * cps, len3p = (len3 + 4 + 3) / 4 * 4;
    auto& tt = ctx->track;
    const size_t n_codes = (size_t)r->n_channels * r->code_len;
    const size_t b_tabs = (size_t)r->n_channels * 4 * len3p * 4, o_codes = (b_tabs + 255) / 256 * 256;
    const bool same = tt.n_ch == r->n_channels && tt.spm == r->spm && tt.period == period &&
                      tt.len3p == len3p && tt.codes.size() == n_codes &&
                      std::equal(tt.codes.begin(), tt.codes.end(), r->codes);
    if (!same) {
        tt.n_ch = 0;
        UGSDR_HIP_TRY(tt.buf.reserve(o_codes + n_codes));
        UGSDR_HIP_TRY(hipMemcpyAsync(tt.buf.as<char>() + o_codes, r->codes, n_codes, hipMemcpyHostToDevice, st));
        const int64_t total = (int64_t)r->n_channels * 4 * len3p;
        const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(track_tables_kernel, dim3(grid), dim3(256), 0, st,
                           reinterpret_cast<const int8_t*>(tt.buf.as<char>() + o_codes), r->code_len, cps, len3,
                           len3p, r->n_channels, tt.buf.as<float>());
        UGSDR_HIP_TRY(hipGetLastError());
        tt.codes.assign(r->codes, r->codes + n_codes);
        tt.n_ch = r->n_channels;
        tt.spm = r->spm;
        tt.period = period;
        tt.len3p = len3p;
    }
    const size_t b_state = (size_t)r->n_channels * sizeof(ugsdr_track_state);
    UGSDR_HIP_TRY(ctx->tables.reserve(b_state));
    char* d = ctx->tables.as<char>();
    UGSDR_HIP_TRY(hipMemcpyAsync(d, ch, b_state, hipMemcpyHostToDevice, st));
    const size_t b_epl = (size_t)r->n_blocks * r->n_channels * 3 * sizeof(float2);
    const size_t b_trace = trace ? (size_t)r->n_blocks * r->n_channels * 4 * sizeof(double) : 0;
    const size_t o_trace = (b_epl + 255) / 256 * 256;
    UGSDR_HIP_TRY(ctx->results.reserve(o_trace + b_trace));
    TrackArgs a{};
    a.src = src;
    a.spm = r->spm;
    a.n_blocks = (int)r->n_blocks;
    a.n_ch = r->n_channels;
    a.fs = r->fs;
    a.spacing = r->spacing_chips;
    a.tabs = tt.buf.as<float>();
    a.len3p = len3p;
    a.state = reinterpret_cast<ugsdr_track_state*>(d);
    a.epl = ctx->results.as<float2>();
    a.trace = trace ? reinterpret_cast<double*>(ctx->results.as<char>() + o_trace) : nullptr;
    ctx->mark(0);
    const int cus = device_cu_count(ctx->device);
    const int rc = r->n_channels <= cus       ? launch_track<kTrackT>(r->fmt, r->n_channels, a, st)
                   : r->n_channels <= 2 * cus ? launch_track<kTrackT / 2>(r->fmt, r->n_channels, a, st)
                                              : launch_track<kTrackT / 4>(r->fmt, r->n_channels, a, st);
    if (rc) return rc;
    UGSDR_HIP_TRY(hipGetLastError());
    ctx->mark(1);
    UGSDR_HIP_TRY(hipMemcpyAsync(epl, a.epl, b_epl, hipMemcpyDeviceToHost, st));
    if (trace) UGSDR_HIP_TRY(hipMemcpyAsync(trace, a.trace, b_trace, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipMemcpyAsync(ch, a.state, b_state, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    if (ctx->profile) {
        // the loop kernel's time goes in the correlation slot
        ctx->last_ms[0] = ctx->last_ms[1] = ctx->last_ms[3] = 0.f;
        UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[2], ctx->ev[0], ctx->ev[1]));
    }
    return UGSDR_OK;
}

}  // namespace ugsdr_hip

extern "C" int ugsdr_hip_track(ugsdr_hip_ctx* ctx, const ugsdr_track_request* req, ugsdr_track_state* ch,
                               float* epl_c32, double* trace) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return ugsdr_hip::track_impl(ctx, req, ch, epl_c32, trace);
}
