// Acquisition surface: FseConfig type list, AcquisitionResult, a generic
// per-operator search engine that runs the reference's ProcessBpsk loop over
// any FseConfig (e.g. the Hip* backends), and the fused HipFastSearchEngine
// that runs the whole PRN x Doppler search in one C-ABI call.
#pragma once

#include <algorithm>
#include <complex>
#include <compare>
#include <cstdint>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "acquisition_result.hpp"
#include "backends.hpp"
#include "context.hpp"

namespace ugsdr {

#ifndef UGSDR_HIP_REFERENCE_BASES
// backend type list with is-base-of gates (src/acquisition/fse.hpp:36-67)
template <auto target_sampling_rate, typename MixerT, typename UpsamplerT, typename MatchedFilterT, typename AbsT,
          typename ReshapeAndSumT, typename MaxIndexT, typename MeanStdDevT, typename ResamplerT>
struct FseConfig {
    constexpr static double acquisition_sampling_rate = target_sampling_rate;
    using MixerType = MixerT;
    using UpsamplerType = UpsamplerT;
    using MatchedFilterType = MatchedFilterT;
    using AbsType = AbsT;
    using ReshapeAndSumType = ReshapeAndSumT;
    using MaxIndexType = MaxIndexT;
    using MeanStdDevType = MeanStdDevT;
    using ResamplerType = ResamplerT;
    static_assert(std::is_base_of_v<Mixer<MixerType>, MixerType>, "Incorrect mixer provided, expected ugsdr::Mixer<T>");
    static_assert(std::is_base_of_v<Upsampler<UpsamplerType>, UpsamplerType>, "Incorrect upsampler provided");
    static_assert(std::is_base_of_v<MatchedFilter<MatchedFilterType>, MatchedFilterType>,
                  "Incorrect matched filter provided");
    static_assert(std::is_base_of_v<Abs<AbsType>, AbsType>, "Incorrect abs provided");
    static_assert(std::is_base_of_v<ReshapeAndSum<ReshapeAndSumType>, ReshapeAndSumType>,
                  "Incorrect reshape_and_sum provided");
    static_assert(std::is_base_of_v<MaxIndex<MaxIndexType>, MaxIndexType>, "Incorrect max_index provided");
    static_assert(std::is_base_of_v<MeanStdDev<MeanStdDevType>, MeanStdDevType>, "Incorrect mean_std_dev provided");
    static_assert(std::is_base_of_v<Resampler<ResamplerType>, ResamplerType>, "Incorrect resampler provided");
};
#endif

// the Hip backend selectable beside ParametricIppFseConfig / ParametricAfFseConfig
template <auto acquisition_sampling_rate>
using ParametricHipFseConfig = FseConfig<acquisition_sampling_rate, HipMixer, HipUpsampler, HipMatchedFilter, HipAbs,
                                         HipReshapeAndSum, HipMaxIndex, HipMeanStdDev, HipResampler>;
using HipFseConfig = ParametricHipFseConfig<8192000>;

// AdjustSamplingRate (src/acquisition/fse.hpp:222-240)
inline double AdjustSamplingRate(double fs, double target) { return ugsdr_adjust_sampling_rate(fs, target); }

// Search parameters of one epoch: the reference's FastSearchEngineBase
// constructor arguments (range, step) plus what its DFE epoch supplies.
struct SearchEpoch {
    const std::complex<float>* samples = nullptr;  // complex float at fs
    std::size_t n_samples = 0;
    double sampling_rate = 0;
    double intermediate_frequency = 0;   // -(central_frequency - 1575.42e6)
};

// ProcessBpsk loop of src/acquisition/fse.hpp:242-276 over any FseConfig:
// every reference operator call goes through Config's backends.
template <typename Config, typename T = float>
class FastSearchEngine {
public:
    FastSearchEngine(double range, double step) : doppler_range(range), doppler_step(step) {}

    // codes: per satellite the +-1 chips of one code period; ms_to_process:
    // coherent length (the reference uses 4)
    std::vector<AcquisitionResult<T>> Process(const SearchEpoch& epoch, const std::vector<Sv>& svs,
                                              const std::vector<std::vector<T>>& codes,
                                              std::size_t ms_to_process = 4) {
        const double fs = epoch.sampling_rate;
        const std::vector<std::complex<T>> signal(epoch.samples, epoch.samples + epoch.n_samples);
        const auto translated = Config::MixerType::Translate(signal, fs, -epoch.intermediate_frequency);
        const double new_fs = AdjustSamplingRate(fs, Config::acquisition_sampling_rate);
        const auto down = Config::ResamplerType::Transform(translated, (std::size_t)new_fs, (std::size_t)fs);
        std::vector<AcquisitionResult<T>> dst;
        for (std::size_t i = 0; i < svs.size(); ++i) {
            const auto repeated = RepeatCodeNTimes(codes[i], ms_to_process);
            const auto code = Config::UpsamplerType::Transform(repeated, (std::size_t)(ms_to_process * new_fs / 1e3));
            const std::vector<std::complex<T>> block(down.begin(), down.begin() + code.size());
            ProcessBpsk(block, code, svs[i], fs, new_fs, epoch.intermediate_frequency, dst);
        }
        std::sort(dst.begin(), dst.end(), [](auto& a, auto& b) { return a.sv_number.id < b.sv_number.id; });
        return dst;
    }

private:
    double doppler_range;
    double doppler_step;
    static constexpr double peak_threshold = 3.3;

    void ProcessBpsk(const std::vector<std::complex<T>>& signal, const std::vector<T>& code, Sv sv, double fs,
                     double new_fs, double intermediate_frequency, std::vector<AcquisitionResult<T>>& dst) {
        AcquisitionResult<T> tmp, max_result;
        const double ratio = fs / new_fs;
        const auto spectrum = Config::MatchedFilterType::PrepareCodeSpectrum(code);
        const auto spm = (std::size_t)(new_fs / 1e3);
        for (double d = -doppler_range; d <= doppler_range; d += doppler_step) {
            const auto translated = Config::MixerType::Translate(signal, new_fs, -d);
            auto matched = Config::MatchedFilterType::FilterOptimized(translated, spectrum);
            const auto abs_value = Config::AbsType::Transform(matched);
            auto peak = Config::ReshapeAndSumType::Transform(abs_value, spm);
            const auto mx = Config::MaxIndexType::Transform(peak);
            const auto ms = Config::MeanStdDevType::Calculate(peak);
            tmp.level = mx.value;
            tmp.sigma = ms.sigma + ms.mean;
            tmp.code_offset = ratio * (double)mx.index;
            tmp.doppler = d + intermediate_frequency;
            if (max_result < tmp) {
                max_result = tmp;
                max_result.output_peak = std::move(peak);
            }
        }
        max_result.intermediate_frequency = intermediate_frequency;
        max_result.sv_number = sv;
        if (max_result.GetSnr() > peak_threshold) dst.push_back(std::move(max_result));
    }
};

struct HipFseOptions {
    double target_sampling_rate = 8192000;  // FseConfig rate
    std::size_t coherent_ms = 4;            // ms_to_process of the reference
    std::size_t noncoherent_blocks = 1;     // K blocks summed after |.|
    bool reshape = true;                    // GetOneMsPeak<reshape, coherent>
    bool coherent = true;
    double threshold = 3.3;
    int precision = UGSDR_PREC_FP32;
    bool output_peak = true;
};

// Fused engine: the whole PRN x Doppler x block search of an epoch in one
// ugsdr_hip_acquire call, replica spectra cached on the device per code.
class HipFastSearchEngine {
public:
    using Options = HipFseOptions;

    HipFastSearchEngine(double range, double step, Options opt = Options(), ugsdr_hip::HipContext* ctx = nullptr)
        : doppler_range(range), doppler_step(step), options(opt),
          context(ctx ? ctx : &ugsdr_hip::HipContext::Default()) {}

    ~HipFastSearchEngine() {
        for (auto& kv : keys) ugsdr_hip_replica_drop(context->get(), kv.second);
    }

    // iq: interleaved samples in `fmt` (UGSDR_FMT_*) at fs
    std::vector<AcquisitionResult<float>> Process(const void* iq, int fmt, std::size_t n_samples, double fs,
                                                  double intermediate_frequency, const std::vector<Sv>& svs,
                                                  const std::vector<std::vector<int8_t>>& codes,
                                                  std::size_t code_ms = 1) {
        return Run(iq, fmt, n_samples, fs, intermediate_frequency, options.target_sampling_rate, options.coherent,
                   svs, codes, code_ms);
    }

    // One band of the digital front end (SignalEpoch::GetSubband with the
    // channel's rate and centre frequency, src/dfe/dfe.hpp); a band the
    // front end does not have is left empty (DigitalFrontend::HasSignal).
    struct Subband {
        const void* iq = nullptr;
        int fmt = UGSDR_FMT_CF32;
        std::size_t n_samples = 0;
        double sampling_rate = 0;
        double central_frequency = 0;
        bool Present() const { return iq != nullptr && n_samples > 0; }
    };

    // FastSearchEngineBase::ProcessGps (src/acquisition/fse.hpp:327-335): GPS
    // L1 C/A (AcquireGoldCodesL1, fse.hpp:278-300) and, only when that finds
    // nothing, L5I and then L5Q (AcquireL5<signal, coherent=false>,
    // fse.hpp:302-325): the IF measured from 1176.45 MHz, the rate from
    // AdjustSamplingRate(fs, 20.46 MHz), the 4 ms correlation folded to 1 ms
    // as complex values before |.|.  svs: the satellites tried (their id is
    // the 0-based PRN); results carry the signal that found them.
    std::vector<AcquisitionResult<float>> ProcessGps(const Subband& l1, const Subband& l5,
                                                     const std::vector<Sv>& svs) {
        std::vector<AcquisitionResult<float>> dst;
        if (l1.Present())
            dst = Acquire(l1, 1575.42e6, options.target_sampling_rate, true, Signal::GpsCoarseAcquisition_L1,
                          ugsdr_code_gps_l1ca, 1023, svs);
        if (!dst.empty() || !l5.Present()) return dst;
        dst = Acquire(l5, 1176.45e6, acquisition_sampling_rate_L5, false, Signal::Gps_L5I, ugsdr_code_gps_l5i,
                      10230, svs);
        if (!dst.empty()) return dst;
        return Acquire(l5, 1176.45e6, acquisition_sampling_rate_L5, false, Signal::Gps_L5Q, ugsdr_code_gps_l5q,
                       10230, svs);
    }

    // FastSearchEngineBase::ProcessGalileo (fse.hpp:356-396): every
    // satellite's 4 ms E1 code (BOC-expanded chips, codes[i]) in the four
    // per-ms sign permutations of fse.hpp:374-379, searched unfolded over
    // 4 ms in one fused call; per satellite the hypothesis of largest level is
    // kept when its SNR clears the threshold.
    std::vector<AcquisitionResult<float>> ProcessGalileo(const void* iq, int fmt, std::size_t n_samples, double fs,
                                                         double intermediate_frequency, const std::vector<Sv>& svs,
                                                         const std::vector<std::vector<int8_t>>& codes) {
        static constexpr int kPermutations[4][4] = {{1, 1, 1, 1}, {1, 1, 1, -1}, {1, 1, -1, -1}, {1, -1, -1, -1}};
        const double new_fs = AdjustSamplingRate(fs, options.target_sampling_rate);
        const auto spm = (std::size_t)(new_fs / 1e3);
        const std::size_t n_fft = 4 * spm;
        std::vector<double> doppler;
        for (double d = -doppler_range; d <= doppler_range; d += doppler_step) doppler.push_back(d);
        std::vector<uint64_t> replica_keys;
        std::vector<float> rep(n_fft);
        for (std::size_t i = 0; i < svs.size(); ++i) {
            const auto& c = codes[i];
            for (int h = 0; h < 4; ++h) {
                if (n_fft % c.size() == 0 && c.size() % 4 == 0) {
                    // whole chips per sample and per ms: the signed chips,
                    // upsampled by the library (the upsampled-code search)
                    std::vector<int8_t> signed_chips(c.size());
                    for (std::size_t k = 0; k < c.size(); ++k)
                        signed_chips[k] = (int8_t)(c[k] * kPermutations[h][k / (c.size() / 4)]);
                    replica_keys.push_back(ChipKey(signed_chips, h, n_fft));
                    continue;
                }
                // Upsampler::Transform (floor index), then ms j times its sign
                for (std::size_t k = 0; k < n_fft; ++k)
                    rep[k] = (float)(c[(k * c.size()) / n_fft] * kPermutations[h][k / spm]);
                replica_keys.push_back(SampleKey(rep, c, h));
            }
        }
        ugsdr_acq_request r{};
        r.iq = iq;
        r.fmt = fmt;
        r.n_samples = (int64_t)n_samples;
        r.fs_in = fs;
        r.fs_acq = new_fs;
        r.mix_frequency = -intermediate_frequency;
        r.intermediate_frequency = intermediate_frequency;
        r.n_fft = (int64_t)n_fft;
        r.n_blocks = 1;
        r.surface_len = (int64_t)n_fft;
        r.reshape = 0;    // ProcessBpsk<false>
        r.coherent = 1;
        r.doppler = doppler.data();
        r.n_doppler = (int64_t)doppler.size();
        r.replica_keys = replica_keys.data();
        r.n_prn = (int64_t)replica_keys.size();
        r.threshold = options.threshold;
        r.precision = options.precision;
        r.hypotheses = 4;
        std::vector<ugsdr_acq_result> out(svs.size());
        ugsdr_hip::check(ugsdr_hip_acquire(context->get(), &r, out.data(), nullptr, nullptr));
        std::vector<AcquisitionResult<float>> dst;
        for (std::size_t i = 0; i < svs.size(); ++i) {
            if (!out[i].detected) continue;
            AcquisitionResult<float> a;
            a.sv_number = svs[i];
            a.doppler = out[i].doppler;
            a.code_offset = out[i].code_offset;
            a.level = out[i].level;
            a.sigma = out[i].sigma;
            a.intermediate_frequency = intermediate_frequency;
            dst.push_back(std::move(a));
        }
        std::sort(dst.begin(), dst.end(), [](auto& a, auto& b) { return a.sv_number.id < b.sv_number.id; });
        return dst;
    }

private:
    static constexpr double acquisition_sampling_rate_L5 = 20.46e6;   // fse.hpp:152

    double doppler_range;
    double doppler_step;
    Options options;

    // AcquireGoldCodesL1 / AcquireL5 of one signal: IF = -(centre - carrier),
    // the mix by -IF, the rate rule and the fused search over every satellite
    std::vector<AcquisitionResult<float>> Acquire(const Subband& band, double carrier, double target_rate,
                                                  bool coherent, Signal signal, int (*code_fn)(int, int8_t*),
                                                  std::size_t code_len, const std::vector<Sv>& svs) {
        std::vector<Sv> tried = svs;
        std::vector<std::vector<int8_t>> codes;
        for (auto& sv : tried) {
            sv.signal = signal;
            std::vector<int8_t> c(code_len);
            ugsdr_hip::check(code_fn(sv.id, c.data()));
            codes.push_back(std::move(c));
        }
        const double intermediate_frequency = -(band.central_frequency - carrier);
        return Run(band.iq, band.fmt, band.n_samples, band.sampling_rate, intermediate_frequency, target_rate,
                   coherent, tried, codes, 1);
    }

    std::vector<AcquisitionResult<float>> Run(const void* iq, int fmt, std::size_t n_samples, double fs,
                                              double intermediate_frequency, double target_rate, bool coherent,
                                              const std::vector<Sv>& svs,
                                              const std::vector<std::vector<int8_t>>& codes, std::size_t code_ms) {
        const double new_fs = AdjustSamplingRate(fs, target_rate);
        const auto spm = (std::size_t)(new_fs / 1e3);
        const std::size_t n_fft = options.coherent_ms * spm;
        std::vector<double> doppler;
        for (double d = -doppler_range; d <= doppler_range; d += doppler_step) doppler.push_back(d);
        std::vector<uint64_t> prn_keys;
        for (std::size_t i = 0; i < svs.size(); ++i) prn_keys.push_back(Key(codes[i], options.coherent_ms / code_ms,
                                                                            n_fft));
        ugsdr_acq_request r{};
        r.iq = iq;
        r.fmt = fmt;
        r.iq_on_device = 0;
        r.n_samples = (int64_t)n_samples;
        r.fs_in = fs;
        r.fs_acq = new_fs;
        r.mix_frequency = -intermediate_frequency;  // Translate(signal, fs, -IF) (fse.hpp:289)
        r.intermediate_frequency = intermediate_frequency;
        r.n_fft = (int64_t)n_fft;
        r.n_blocks = (int64_t)options.noncoherent_blocks;
        r.surface_len = (int64_t)(options.reshape ? spm : n_fft);
        r.reshape = options.reshape;
        r.coherent = coherent;
        r.doppler = doppler.data();
        r.n_doppler = (int64_t)doppler.size();
        r.replica_keys = prn_keys.data();
        r.n_prn = (int64_t)prn_keys.size();
        r.threshold = options.threshold;
        r.precision = options.precision;
        std::vector<ugsdr_acq_result> out(svs.size());
        std::vector<float> peaks(options.output_peak ? svs.size() * (std::size_t)r.surface_len : 0);
        ugsdr_hip::check(ugsdr_hip_acquire(context->get(), &r, out.data(), nullptr,
                                           options.output_peak ? peaks.data() : nullptr));
        std::vector<AcquisitionResult<float>> dst;
        for (std::size_t i = 0; i < svs.size(); ++i) {
            if (!out[i].detected) continue;
            AcquisitionResult<float> a;
            a.sv_number = svs[i];
            a.doppler = out[i].doppler;
            a.code_offset = out[i].code_offset;
            a.level = out[i].level;
            a.sigma = out[i].sigma;
            a.intermediate_frequency = intermediate_frequency;
            if (options.output_peak)
                a.output_peak.assign(peaks.begin() + i * r.surface_len, peaks.begin() + (i + 1) * r.surface_len);
            dst.push_back(std::move(a));
        }
        std::sort(dst.begin(), dst.end(), [](auto& a, auto& b) { return a.sv_number.id < b.sv_number.id; });
        return dst;
    }
    ugsdr_hip::HipContext* context;
    std::vector<std::pair<std::vector<int64_t>, uint64_t>> keys;

    // replica from samples, cached by (code, hypothesis, length, precision)
    uint64_t SampleKey(const std::vector<float>& samples, const std::vector<int8_t>& code, int hypothesis) {
        std::vector<int64_t> id{-1 - hypothesis, (int64_t)samples.size(), options.precision};
        id.insert(id.end(), code.begin(), code.end());
        for (auto& kv : keys)
            if (kv.first == id) return kv.second;
        uint64_t k = 0;
        ugsdr_hip::check(ugsdr_hip_replica_put_samples(context->get(), samples.data(), (int64_t)samples.size(),
                                                       options.precision, &k));
        keys.emplace_back(std::move(id), k);
        return k;
    }

    // signed chips of hypothesis h, cached by (hypothesis, length, precision, chips)
    uint64_t ChipKey(const std::vector<int8_t>& chips, int hypothesis, std::size_t n_fft) {
        std::vector<int64_t> id{-5 - hypothesis, (int64_t)n_fft, options.precision};
        id.insert(id.end(), chips.begin(), chips.end());
        for (auto& kv : keys)
            if (kv.first == id) return kv.second;
        uint64_t k = 0;
        ugsdr_hip::check(ugsdr_hip_replica_put(context->get(), chips.data(), (int64_t)chips.size(), 1,
                                               (int64_t)n_fft, options.precision, &k));
        keys.emplace_back(std::move(id), k);
        return k;
    }

    uint64_t Key(const std::vector<int8_t>& code, std::size_t repeats, std::size_t n_fft) {
        std::vector<int64_t> id{(int64_t)repeats, (int64_t)n_fft, options.precision};
        id.insert(id.end(), code.begin(), code.end());
        for (auto& kv : keys)
            if (kv.first == id) return kv.second;
        uint64_t k = 0;
        ugsdr_hip::check(ugsdr_hip_replica_put(context->get(), code.data(), (int64_t)code.size(), (int64_t)repeats,
                                               (int64_t)n_fft, options.precision, &k));
        keys.emplace_back(std::move(id), k);
        return k;
    }
};

}  // namespace ugsdr
