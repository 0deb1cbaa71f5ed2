// GPU-resident tracking loops (Tracker::Track, src/tracking/tracker.hpp:156-197).
//
// One workgroup per channel walks the epochs in order.  Per 1 ms epoch:
//   - thread 0 plans it: the code phase of the epoch (TrackSingleSatellite,
//     tracker.hpp:165), the early / prompt / late split points
//     (CorrelateSplit, tracking_parameters.hpp:471-486) and the carrier tone
//     of the in-place Mixer::Translate (mixer.hpp:33-41), then advances the
//     carrier phase (UpdatePhase, tracking_parameters.hpp:163-165);
//   - every thread mixes its samples and accumulates the six partial
//     correlations against the 3-period replica table;
//   - thread 0 combines them (AddWithPhase, tracking_parameters.hpp:89-98)
//     and runs the PLL/FLL and DLL updates (:495-528) in FP64, so the next
//     epoch uses the new state without a trip to the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "formats.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

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
    const float* code3;          // n_ch x code3_len replica tables
    int64_t code3_len;
    ugsdr_track_state* state;    // n_ch
    float2* epl;                 // n_blocks x n_ch x 3
    double* trace;               // n_blocks x n_ch x 4, or null
};

struct EpochPlan {
    int first_len[3], first_phase[3], second_phase[3];
    double rel[3];
    double df, ph_turns;
    int mix;
};

// TrackingParameters::Pll (tracking_parameters.hpp:495-511).  The prompt is
// complex<float>: the I/Q ratio and the cross / dot products are float, the
// discriminators and the state FP64.
__device__ __forceinline__ void pll_update(ugsdr_track_state& s, float2 p) {
    const double e = p.x != 0.f ? atan((double)(p.y / p.x)) / (kPi * 2.0) : 0.0;
    const float qx = s.previous_prompt[0], qy = s.previous_prompt[1];
    const float cross = p.x * qy - qx * p.y;
    const float dot = fabsf(p.x * qx + qy * p.y);
    const double ferr = (double)atan2f(cross, dot) / kPi;
    s.carrier_frequency += gains::kPll1 * e - gains::kPll2 * s.carrier_phase_error - gains::kPll3 * ferr;
    s.carrier_phase_error = e;
    s.carrier_phase += e * kPi / 2;
    s.previous_prompt[0] = p.x;
    s.previous_prompt[1] = p.y;
}

// TrackingParameters::Dll (tracking_parameters.hpp:513-528): normalised
// early-minus-late envelope in float, NCO and code phase in FP64
__device__ __forceinline__ void dll_update(ugsdr_track_state& s, float2 e, float2 l, double fs) {
    const float ea = hypotf(e.x, e.y), la = hypotf(l.x, l.y);
    double err = 0.0;
    if (ea + la != 0.f) err = (ea - la) / (ea + la);
    s.code_nco += (gains::kDll1 * (err - s.code_error) + gains::kDll2 * err) * (s.base_code_frequency / 1.023e6);
    s.code_error = err;
    s.code_frequency = s.base_code_frequency - s.code_nco;
    s.code_phase -= fs / 1000 / 2 * (s.code_frequency / s.base_code_frequency - 1);
}

__global__ __launch_bounds__(kTrackT) void track_kernel(TrackArgs a) {
    const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ EpochPlan plan;
    __shared__ float2 red[(kTrackT / 64) * 6];
    __shared__ ugsdr_track_state s;
    const float* code3 = a.code3 + (int64_t)c * a.code3_len;
    if (tid == 0) s = a.state[c];
    for (int b = 0; b < a.n_blocks; ++b) {
        if (tid == 0) {
            const int64_t spm = a.spm;
            const double cps = (double)spm * s.code_period_ms;
            // TrackSingleSatellite (tracker.hpp:165): the code phase of this epoch
            const double cp = s.code_phase - a.fs / 1e3 * (double)(s.epochs % s.code_period_ms);
            const double off = a.fs / s.base_code_frequency * a.spacing;  // GetSamplesPerChip() * spacing
            const double taps[3] = {cp + off, cp, cp - off};
            for (int t = 0; t < 3; ++t) {
                double p = taps[t];
                while (p < 0) p += cps;
                if (p > cps) p = fmod(p, cps);
                plan.first_len[t] = (int)((int64_t)ceil(p) % spm);
                plan.first_phase[t] = (int)(int64_t)ceil(2 * cps - p);
                plan.second_phase[t] = plan.first_phase[t] + plan.first_len[t];
                plan.rel[t] = fmod(p, (double)spm) / (double)spm;
            }
            // Translate(x, fs, -f_carr, -phase_carr), in place
            double phase = fmod(-s.carrier_phase, 2.0 * kPi);
            if (phase < 0) phase += 2.0 * kPi;
            double f = -s.carrier_frequency;
            plan.mix = f != 0.0;
            if (f < 0) f = a.fs + f;
            plan.df = f / a.fs;
            plan.ph_turns = phase / (2.0 * kPi);
            s.carrier_phase += 2 * kPi * s.carrier_frequency / 1000.0;  // UpdatePhase
        }
        __syncthreads();
        float2 acc[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) acc[i] = make_float2(0.f, 0.f);
        const int64_t base = (int64_t)b * a.spm;
        const double df = plan.df, pt = plan.ph_turns;
        const int mix = plan.mix;
        int fl[3], fp[3], sp[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            fl[t] = plan.first_len[t];
            fp[t] = plan.first_phase[t];
            sp[t] = plan.second_phase[t];
        }
        for (int n = tid; n < a.spm; n += kTrackT) {
            float2 x = load_sample(a.src, base + n);
            if (mix) {
                double ph = fma((double)n, df, pt);
                ph -= rint(ph);
                float sn, cs;
                sincosf((float)(2.0 * kPi * ph), &sn, &cs);
                x = cmul(x, make_float2(cs, sn));
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const bool first = n < fl[t];
                const float cv = code3[first ? fp[t] + n : sp[t] + (n - fl[t])];
                float2& ac = acc[2 * t + (first ? 0 : 1)];
                ac.x = fmaf(x.x, cv, ac.x);
                ac.y = fmaf(x.y, cv, ac.y);
            }
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            acc[i].x = wave_sum(acc[i].x);
            acc[i].y = wave_sum(acc[i].y);
            if (lane == 0) red[wave * 6 + i] = acc[i];
        }
        __syncthreads();
        if (tid == 0) {
            float2 part[6];
            for (int i = 0; i < 6; ++i) {
                part[i] = make_float2(0.f, 0.f);
                for (int w = 0; w < kTrackT / 64; ++w) part[i] = cadd(part[i], red[w * 6 + i]);
            }
            float2 epl[3];
            for (int t = 0; t < 3; ++t) {
                float2 lhs = part[2 * t], rhs = part[2 * t + 1];
                if (plan.rel[t] < 0.5) {
                    const float2 tmp = lhs;
                    lhs = rhs;
                    rhs = tmp;
                }
                const double wl = (double)lhs.x / plan.rel[t], wr = (double)rhs.x / (1.0 - plan.rel[t]);
                epl[t] = fabs(wl + wr) > fabs(wl) ? cadd(lhs, rhs) : csub(lhs, rhs);
            }
            float2* o = a.epl + ((int64_t)b * a.n_ch + c) * 3;
            o[0] = epl[0];
            o[1] = epl[1];
            o[2] = epl[2];
            pll_update(s, epl[1]);
            dll_update(s, epl[0], epl[2], a.fs);
            ++s.epochs;
            if (a.trace) {
                double* tr = a.trace + ((int64_t)b * a.n_ch + c) * 4;
                tr[0] = -s.carrier_phase / (2 * kPi);
                tr[1] = s.carrier_frequency;
                tr[2] = s.code_phase;
                tr[3] = s.code_frequency;
            }
        }
        __syncthreads();
    }
    if (tid == 0) a.state[c] = s;
}

static int track_impl(ugsdr_hip_ctx* ctx, const ugsdr_track_request* r, ugsdr_track_state* ch, float* epl,
                      double* trace) {
    UGSDR_CHECK_ARG(r && ch && epl, "null argument");
    UGSDR_CHECK_ARG(r->iq && r->spm > 0 && r->n_blocks > 0 && r->fs > 0, "bad signal");
    UGSDR_CHECK_ARG(r->n_channels > 0 && r->codes && r->code_len > 0, "bad channels");
    UGSDR_CHECK_ARG(r->spacing_chips >= 0 && r->spacing_chips < 1, "bad early/late spacing");
    UGSDR_CHECK_ARG(r->n_blocks < (1 << 30) && r->spm < (1 << 24) && r->spm * r->n_blocks < ((int64_t)1 << 31),
                    "at most 2^31 - 1 samples a call");
    const int period = ch[0].code_period_ms;
    for (int c = 0; c < r->n_channels; ++c) {
        UGSDR_CHECK_ARG(ch[c].code_period_ms == period && period >= 1, "channels of one call share a code period");
        UGSDR_CHECK_ARG(ch[c].base_code_frequency > 0, "bad code frequency");
    }
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
    // 3-period replica tables at fs (Codes::FillCodesImpl, tracker.hpp:39-51):
    // floor-index upsampling of the chips repeated three times
    const int64_t cps = r->spm * period, len3 = 3 * cps, chips3 = 3 * r->code_len;
    std::vector<float> code3((size_t)r->n_channels * len3);
    for (int c = 0; c < r->n_channels; ++c)
        for (int64_t i = 0; i < len3; ++i)
            code3[(size_t)c * len3 + i] =
                (float)r->codes[(size_t)c * r->code_len + ((i * chips3 / len3) % chips3) % r->code_len];
    const size_t b_code = code3.size() * 4, b_state = (size_t)r->n_channels * sizeof(ugsdr_track_state);
    const size_t o_state = (b_code + 255) / 256 * 256;
    UGSDR_HIP_TRY(ctx->tables.reserve(o_state + b_state));
    char* d = ctx->tables.as<char>();
    UGSDR_HIP_TRY(hipMemcpyAsync(d, code3.data(), b_code, hipMemcpyHostToDevice, st));
    UGSDR_HIP_TRY(hipMemcpyAsync(d + o_state, ch, b_state, hipMemcpyHostToDevice, st));
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
    a.code3 = reinterpret_cast<const float*>(d);
    a.code3_len = len3;
    a.state = reinterpret_cast<ugsdr_track_state*>(d + o_state);
    a.epl = ctx->results.as<float2>();
    a.trace = trace ? reinterpret_cast<double*>(ctx->results.as<char>() + o_trace) : nullptr;
    hipLaunchKernelGGL(track_kernel, dim3((unsigned)r->n_channels), dim3(kTrackT), 0, st, a);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(epl, a.epl, b_epl, hipMemcpyDeviceToHost, st));
    if (trace) UGSDR_HIP_TRY(hipMemcpyAsync(trace, a.trace, b_trace, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipMemcpyAsync(ch, a.state, b_state, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
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
