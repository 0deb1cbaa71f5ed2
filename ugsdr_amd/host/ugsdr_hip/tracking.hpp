// Tracking-side surface: the TrackingParametersConfig type list, the
// CorrelateSplit / AddWithPhase tap of the reference written over any
// Config::CorrelatorType, HipCorrelatorBank, which evaluates every channel's
// taps over many 1 ms blocks in one C-ABI call, and HipTracker, the closed
// tracking loops of every channel run on the device.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <span>
#include <utility>
#include <vector>

#include "backends.hpp"
#include "context.hpp"
#include "acquisition_result.hpp"

namespace ugsdr {

// src/tracking/tracking_parameters.hpp:15-37
template <typename AbsT, typename CorrelatorT, typename MatchedFilterT, typename MixerT, typename ReshapeAndSumT,
          typename UpsamplerT>
struct TrackingParametersConfig {
    using AbsType = AbsT;
    using CorrelatorType = CorrelatorT;
    using MatchedFilterType = MatchedFilterT;
    using MixerType = MixerT;
    using ReshapeAndSumType = ReshapeAndSumT;
    using UpsamplerType = UpsamplerT;
    static_assert(std::is_base_of_v<Correlator<CorrelatorType>, CorrelatorType>, "Incorrect correlator provided");
    static_assert(std::is_base_of_v<Mixer<MixerType>, MixerType>, "Incorrect mixer provided");
};

using HipTrackingParametersConfig =
    TrackingParametersConfig<HipAbs, HipCorrelator, HipMatchedFilter, HipMixer, HipReshapeAndSum, HipUpsampler>;

// AddWithPhase (src/tracking/tracking_parameters.hpp:89-98)
template <typename T>
std::complex<T> AddWithPhase(std::complex<T> lhs, std::complex<T> rhs, double relative_phase) {
    if (relative_phase < 0.5) std::swap(lhs, rhs);
    const double wl = lhs.real() / relative_phase;
    const double wr = rhs.real() / (1.0 - relative_phase);
    return std::abs(wl + wr) > std::abs(wl) ? lhs + rhs : lhs - rhs;
}

// CorrelateSplit (src/tracking/tracking_parameters.hpp:471-493) over Config's correlator
template <typename Config, typename T>
std::complex<T> CorrelateSplit(const std::vector<std::complex<T>>& translated, const std::vector<T>& full_code,
                               double sampling_rate, double code_period_ms, double phase) {
    const auto spm = (std::size_t)(sampling_rate / 1e3);
    const double cps = (double)spm * code_period_ms;
    while (phase < 0) phase += cps;
    if (phase > cps) phase = std::fmod(phase, cps);
    const auto first_len = (std::size_t)std::ceil(phase) % spm;
    const auto second_len = spm - first_len;
    const auto first_phase = (std::size_t)std::ceil(2 * cps - phase);
    const auto second_phase = first_phase + first_len;
    auto first = Config::CorrelatorType::Correlate(std::span<const std::complex<T>>(translated.data(), first_len),
                                                   std::span<const T>(full_code.data() + first_phase, first_len));
    auto second = Config::CorrelatorType::Correlate(
        std::span<const std::complex<T>>(translated.data() + first_len, second_len),
        std::span<const T>(full_code.data() + second_phase, second_len));
    return AddWithPhase(first, second, std::fmod(phase, (double)spm) / (double)spm);
}

// Correlator bank over n_blocks consecutive blocks for many channels at once.
class HipCorrelatorBank {
public:
    using Channel = ugsdr_channel_state;

    HipCorrelatorBank(std::vector<double> taps_chips, ugsdr_hip::HipContext* ctx = nullptr)
        : taps(std::move(taps_chips)), context(ctx ? ctx : &ugsdr_hip::HipContext::Default()) {}

    // NCO mode: FP64 carrier / code NCOs per sample; channels advance in place.
    // Returns n_blocks x n_channels x n_taps correlations.
    std::vector<std::complex<float>> CorrelateNco(const void* iq, int fmt, std::size_t spm, std::size_t n_blocks,
                                                  double fs, std::vector<Channel>& channels,
                                                  const std::vector<int8_t>& codes, std::size_t code_len) {
        return Run(iq, fmt, spm, n_blocks, fs, UGSDR_BANK_NCO, channels, codes, code_len, 0, 1);
    }

    // Tracker::GetEpl semantics (src/tracking/tracker.hpp:119-139)
    std::vector<std::complex<float>> CorrelateSplit(const void* iq, int fmt, std::size_t spm, std::size_t n_blocks,
                                                    double fs, std::vector<Channel>& channels,
                                                    const std::vector<int8_t>& codes, std::size_t code_len,
                                                    double samples_per_chip, double code_period_ms = 1.0) {
        return Run(iq, fmt, spm, n_blocks, fs, UGSDR_BANK_REF_SPLIT, channels, codes, code_len, samples_per_chip,
                   code_period_ms);
    }

private:
    std::vector<double> taps;
    ugsdr_hip::HipContext* context;

    std::vector<std::complex<float>> Run(const void* iq, int fmt, std::size_t spm, std::size_t n_blocks, double fs,
                                         int mode, std::vector<Channel>& channels, const std::vector<int8_t>& codes,
                                         std::size_t code_len, double spc, double period_ms) {
        ugsdr_bank_request r{};
        r.iq = iq;
        r.fmt = fmt;
        r.spm = (int64_t)spm;
        r.n_blocks = (int64_t)n_blocks;
        r.fs = fs;
        r.mode = mode;
        r.n_channels = (int32_t)channels.size();
        r.n_taps = (int32_t)taps.size();
        r.taps_chips = taps.data();
        r.codes = codes.data();
        r.code_len = (int64_t)code_len;
        r.samples_per_chip = spc;
        r.code_period_ms = period_ms;
        std::vector<std::complex<float>> out(n_blocks * channels.size() * taps.size());
        ugsdr_hip::check(ugsdr_hip_correlate_bank(context->get(), &r, channels.data(),
                                                  reinterpret_cast<float*>(out.data())));
        return out;
    }
};

// Tracker (src/tracking/tracker.hpp:99-197) for the channels of one signal at
// one sampling rate: the loops (PLL/FLL and DLL of TrackingParameters) run on
// the device, and each channel keeps the per-epoch records TrackingParameters
// keeps (early / prompt / late, phases, frequencies, code phases, code
// frequencies).
class HipTracker {
public:
    struct Channel {
        Sv sv;
        ugsdr_track_state state{};
        std::vector<std::complex<float>> early, prompt, late;
        std::vector<double> phases, frequencies, code_phases, code_frequencies;
    };

    // TrackingParameters(acquisition, ...) (tracking_parameters.hpp:118-140)
    // for signals whose code period is code_period_ms: code phase from the
    // acquisition's code offset, carrier from its Doppler.
    HipTracker(double sampling_rate, const std::vector<AcquisitionResult<float>>& acquisitions,
               const std::vector<std::vector<int8_t>>& codes, double base_code_frequency = 1.023e6,
               int code_period_ms = 1, ugsdr_hip::HipContext& ctx = ugsdr_hip::HipContext::Default())
        : fs(sampling_rate), context(&ctx), chips(codes) {
        for (const auto& a : acquisitions) {
            Channel c;
            c.sv = a.sv_number;
            c.state.code_phase = a.code_offset;
            c.state.carrier_frequency = a.doppler;
            c.state.code_frequency = c.state.base_code_frequency = base_code_frequency;
            c.state.code_period_ms = code_period_ms;
            channels.push_back(std::move(c));
        }
    }

    // Tracker::Track over n_epochs 1 ms epochs of interleaved samples in fmt
    // at the tracker's rate (host, or device memory with on_device)
    void Track(const void* samples, int fmt, std::size_t n_epochs, bool on_device = false, double spacing = 0.25) {
        const std::size_t n = channels.size();
        if (!n) return;
        const auto spm = (int64_t)(fs / 1e3);
        std::vector<ugsdr_track_state> st(n);
        std::vector<int8_t> codes;
        for (std::size_t i = 0; i < n; ++i) {
            st[i] = channels[i].state;
            codes.insert(codes.end(), chips[i].begin(), chips[i].end());
        }
        ugsdr_track_request r{};
        r.iq = samples;
        r.fmt = fmt;
        r.iq_on_device = on_device ? 1 : 0;
        r.spm = spm;
        r.n_blocks = (int64_t)n_epochs;
        r.fs = fs;
        r.n_channels = (int32_t)n;
        r.codes = codes.data();
        r.code_len = (int64_t)chips[0].size();
        r.spacing_chips = spacing;
        std::vector<std::complex<float>> epl(n_epochs * n * 3);
        std::vector<double> trace(n_epochs * n * 4);
        ugsdr_hip::check(ugsdr_hip_track(context->get(), &r, st.data(), reinterpret_cast<float*>(epl.data()),
                                         trace.data()));
        for (std::size_t i = 0; i < n; ++i) {
            Channel& c = channels[i];
            c.state = st[i];
            for (std::size_t b = 0; b < n_epochs; ++b) {
                const auto* e = &epl[(b * n + i) * 3];
                const auto* t = &trace[(b * n + i) * 4];
                c.early.push_back(e[0]);
                c.prompt.push_back(e[1]);
                c.late.push_back(e[2]);
                c.phases.push_back(t[0]);
                c.frequencies.push_back(t[1]);
                c.code_phases.push_back(t[2]);
                c.code_frequencies.push_back(t[3]);
            }
        }
    }

    const std::vector<Channel>& GetTrackingParameters() const { return channels; }

private:
    double fs;
    ugsdr_hip::HipContext* context;
    std::vector<std::vector<int8_t>> chips;
    std::vector<Channel> channels;
};

}  // namespace ugsdr
