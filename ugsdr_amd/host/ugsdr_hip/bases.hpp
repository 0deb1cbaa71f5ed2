// CRTP operator bases with the reference's public surface (names, argument
// meaning, error behaviour), so the Hip* backends compile and run on their
// own.  Inside a reference build, define UGSDR_HIP_REFERENCE_BASES and include
// the reference's own bases first: the Hip* backends then derive from those.
//
//   Mixer<Impl>          src/mixer/mixer.hpp:13-69
//   MatchedFilter<Impl>  src/matched_filter/matched_filter.hpp:11-43
//   Abs<Impl>            src/math/abs.hpp:9-19
//   ReshapeAndSum<Impl>  src/math/reshape_and_sum.hpp:11-32
//   MaxIndex<Impl>       src/math/max_index.hpp:11-25
//   MeanStdDev<Impl>     src/math/mean_stddev.hpp:9-25
//   Resampler<Impl>      src/resample/resampler.hpp:11-26
//   Upsampler<Impl>      src/resample/upsampler.hpp:9-24
//   Correlator<Impl>     src/correlator/correlator.hpp:11-21
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <numbers>
#include <span>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace ugsdr {

#ifndef UGSDR_HIP_REFERENCE_BASES

template <typename T>
struct IsContainer : std::false_type {};
template <typename T>
struct IsContainer<std::vector<T>> : std::true_type {};
template <typename T>
struct IsContainer<std::span<T>> : std::true_type {};

template <typename T>
struct IsComplexContainer : std::false_type {};
template <typename T>
struct IsComplexContainer<std::vector<std::complex<T>>> : std::true_type {};
template <typename T>
struct IsComplexContainer<std::span<std::complex<T>>> : std::true_type {};

template <typename T>
concept Container = IsContainer<T>::value;
template <typename T>
concept ComplexContainer = IsComplexContainer<T>::value;

// RepeatCodeNTimes (src/common.hpp:78-85)
template <typename T>
auto RepeatCodeNTimes(std::vector<T> code, std::size_t repeats) {
    const auto period = code.size();
    code.reserve(period * repeats);
    for (std::size_t i = 1; i < repeats; ++i)
        for (std::size_t j = 0; j < period; ++j) code.push_back(code[j]);
    return code;
}

template <typename MixerImpl>
class Mixer {
protected:
    double sampling_rate = 0.0;
    double mixer_frequency = 0.0;
    double mixer_phase = 0.0;

    template <typename T>
    static void FixPhase(T& phase) {
        phase = std::fmod(phase, 2 * std::numbers::pi_v<T>);
        if (phase < 0) phase += 2 * std::numbers::pi_v<T>;
    }

public:
    Mixer(double sampling_freq, double frequency, double phase)
        : sampling_rate(sampling_freq), mixer_frequency(frequency), mixer_phase(phase) {}

    // in place: no-op for a zero frequency
    template <ComplexContainer T>
    static void Translate(T& src_dst, double sampling_freq, double frequency, double phase = 0) {
        FixPhase(phase);
        if (std::abs(frequency) == 0.0) return;
        if (frequency < 0) frequency = sampling_freq + frequency;
        MixerImpl::Process(src_dst, sampling_freq, frequency, phase);
    }

    template <ComplexContainer T>
    static auto Translate(const T& src, double sampling_freq, double frequency, double phase = 0) {
        if (frequency < 0) frequency = sampling_freq + frequency;
        FixPhase(phase);
        return MixerImpl::Process(src, sampling_freq, frequency, phase);
    }

    // stateful: keeps phase continuity across consecutive 1 ms calls
    template <ComplexContainer T>
    void Translate(T& src_dst) {
        Translate(src_dst, sampling_rate, mixer_frequency, mixer_phase);
        auto frac = std::fmod(mixer_frequency / 1000.0, 1.0);
        if (frac < 0) frac += 1;
        mixer_phase += 2 * std::numbers::pi_v<double> * frac;
    }

    auto GetFrequency() const { return mixer_frequency; }
    void SetPhase(double phase) { mixer_phase = phase; }
};

template <typename FilterImpl>
class MatchedFilter {
public:
    template <ComplexContainer T1, Container T2>
    static void Filter(T1& src_dst, const T2& impulse_response) {
        FilterImpl::Process(src_dst, impulse_response);
    }
    template <ComplexContainer T1, Container T2>
    static auto Filter(const T1& src, const T2& impulse_response) {
        auto dst = src;
        FilterImpl::Process(dst, impulse_response);
        return dst;
    }
    template <Container T>
    static auto PrepareCodeSpectrum(const T& impulse_response) {
        return FilterImpl::Prepare(impulse_response);
    }
    template <ComplexContainer T1, Container T2>
    static void FilterOptimized(T1& src_dst, const T2& spectrum) {
        FilterImpl::ProcessOptimized(src_dst, spectrum);
    }
    template <ComplexContainer T1, Container T2>
    static auto FilterOptimized(const T1& src, const T2& spectrum) {
        auto dst = src;
        FilterImpl::ProcessOptimized(dst, spectrum);
        return dst;
    }
};

// ShortTimeFourierTransform (src/math/stft.hpp:13-29): columns of a signal,
// or the overlap-add of columns back into a signal
template <typename StftImpl>
class ShortTimeFourierTransform {
public:
    template <ComplexContainer T>
    static auto Transform(const T& src, std::size_t window_length = 128, std::size_t overlap_length = 96) {
        return StftImpl::Process(src, window_length, overlap_length);
    }
    template <Container T>
    requires ComplexContainer<typename T::value_type>
    static auto Transform(const T& src, std::size_t window_length = 128, std::size_t overlap_length = 96) {
        return StftImpl::Process(src, window_length, overlap_length);
    }
};

template <typename AbsImpl>
class Abs {
public:
    template <ComplexContainer T>
    [[nodiscard]] static auto Transform(const T& src) {
        return AbsImpl::Process(src);
    }
};

template <typename SumImpl>
class ReshapeAndSum {
protected:
    template <Container T>
    static void CheckBlockSize(const T& vec, std::size_t block_size) {
        if (block_size == 0 || vec.size() % block_size != 0)
            throw std::runtime_error("Signal size is not divided by block size");
    }

public:
    template <Container T>
    static void Transform(T& src_dst, std::size_t block_size) {
        CheckBlockSize(src_dst, block_size);
        SumImpl::Process(src_dst, block_size);
    }
    template <Container T>
    static auto Transform(const T& src, std::size_t block_size) {
        CheckBlockSize(src, block_size);
        return SumImpl::Process(src, block_size);
    }
};

template <typename MaxIndexImpl>
class MaxIndex {
public:
    template <typename T, typename Tsz = std::size_t>
    struct Result {
        T value{};
        Tsz index{};
    };
    template <Container T>
    static auto Transform(const T& src) {
        return MaxIndexImpl::Process(src);
    }
};

template <typename StdDevImpl>
class MeanStdDev {
public:
    template <typename T>
    struct Result {
        T mean{};
        T sigma{};
    };
    template <Container T>
    static auto Calculate(const T& src) {
        return StdDevImpl::Process(src);
    }
};

template <typename ResamplerImpl>
class Resampler {
public:
    template <Container T>
    static void Transform(T& src_dst, std::size_t new_sampling_rate, std::size_t old_sampling_rate) {
        ResamplerImpl::Process(src_dst, new_sampling_rate, old_sampling_rate);
    }
    template <Container T>
    static auto Transform(const T& src, std::size_t new_sampling_rate, std::size_t old_sampling_rate) {
        auto dst = src;
        ResamplerImpl::Process(dst, new_sampling_rate, old_sampling_rate);
        return dst;
    }
};

template <typename UpsamplerImpl>
class Upsampler {
public:
    template <Container T>
    static void Transform(T& src_dst, std::size_t samples) {
        UpsamplerImpl::Process(src_dst, samples);
    }
    template <Container T>
    static auto Transform(const T& src, std::size_t samples) {
        auto dst = src;
        UpsamplerImpl::Process(dst, samples);
        return dst;
    }
};

template <typename CorrelatorImpl>
class Correlator {
public:
    template <Container T1, Container T2>
    [[nodiscard]] static auto Correlate(const T1& signal, const T2& code) {
        return CorrelatorImpl::Process(signal, code);
    }
};

#endif  // UGSDR_HIP_REFERENCE_BASES

}  // namespace ugsdr
