// Hip* operator backends: drop-ins beside the reference's Sequential*, Ipp*
// and Af* implementations.  Each Process hook has the shape the CRTP base
// calls and forwards one operator call to the C ABI (host data in and out,
// like the reference's ArrayFire proxies, src/helpers/af_array_proxy.hpp).
#pragma once

#include <complex>
#include <cstddef>
#include <span>
#include <type_traits>
#include <vector>

#include "bases.hpp"
#include "context.hpp"

namespace ugsdr {

namespace hip_detail {
inline ugsdr_hip_ctx* ctx() { return ugsdr_hip::HipContext::Default().get(); }
inline const float* cf(const std::complex<float>* p) { return reinterpret_cast<const float*>(p); }
inline float* cf(std::complex<float>* p) { return reinterpret_cast<float*>(p); }

// The Hip* backends compute in FP32 (complex<float>), the reference's float
// path.  The reference's typed tests also instantiate its operators over
// double (test/basic_tests.cpp); those instantiations have no device
// counterpart here and stop at compile time with this message instead of
// narrowing to float in silence.
template <typename T>
constexpr bool is_fp32 = std::is_same_v<T, float>;
template <typename T>
void fp32_only() {
    static_assert(is_fp32<T>, "ugsdr Hip* backends compute in FP32 (std::complex<float>); "
                              "instantiations over double have no device counterpart");
}
}  // namespace hip_detail

// IppMixer::Process counterpart (src/mixer/ipp_mixer.hpp:47-88); frequency is
// already folded to [0, fs) and the phase fixed by the base.
class HipMixer : public Mixer<HipMixer> {
protected:
    friend class Mixer<HipMixer>;

    static void Process(std::vector<std::complex<float>>& src_dst, double fs, double f, double phase = 0) {
        ugsdr_hip::check(ugsdr_hip_mix(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                       hip_detail::cf(src_dst.data()), (int64_t)src_dst.size(), fs, f, phase));
    }
    static void Process(std::span<std::complex<float>> src_dst, double fs, double f, double phase = 0) {
        ugsdr_hip::check(ugsdr_hip_mix(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                       hip_detail::cf(src_dst.data()), (int64_t)src_dst.size(), fs, f, phase));
    }
    static auto Process(const std::vector<std::complex<float>>& src, double fs, double f, double phase = 0) {
        std::vector<std::complex<float>> dst(src.size());
        ugsdr_hip::check(ugsdr_hip_mix(hip_detail::ctx(), hip_detail::cf(src.data()), hip_detail::cf(dst.data()),
                                       (int64_t)src.size(), fs, f, phase));
        return dst;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<std::complex<U>>&, double, double, double = 0) {
        hip_detail::fp32_only<U>();
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static std::vector<std::complex<U>> Process(const std::vector<std::complex<U>>& src, double, double, double = 0) {
        hip_detail::fp32_only<U>();
        return src;
    }

public:
    HipMixer(double sampling_freq, double frequency, double phase) : Mixer<HipMixer>(sampling_freq, frequency, phase) {}
};

// TableMixer counterpart (src/mixer/table_mixer.hpp:27-105), the mixer of the
// reference's non-IPP tracking and DFE configurations: 32-bit NCO and the
// 80-entry float sine table, bit-exact with the reference.  Like TableMixer it
// has only the in-place Process, so only Mixer's in-place and member
// Translate overloads instantiate.
class HipTableMixer : public Mixer<HipTableMixer> {
protected:
    friend class Mixer<HipTableMixer>;

    static void Process(std::vector<std::complex<float>>& src_dst, double fs, double f, double phase = 0) {
        ugsdr_hip::check(ugsdr_hip_table_mix(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                             (int64_t)src_dst.size(), fs, f, phase, 0));
    }
    static void Process(std::span<std::complex<float>> src_dst, double fs, double f, double phase = 0) {
        ugsdr_hip::check(ugsdr_hip_table_mix(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                             (int64_t)src_dst.size(), fs, f, phase, 0));
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<std::complex<U>>&, double, double, double = 0) {
        hip_detail::fp32_only<U>();
    }

public:
    HipTableMixer(double sampling_freq, double frequency, double phase)
        : Mixer<HipTableMixer>(sampling_freq, frequency, phase) {}
};

// IppMatchedFilter counterpart (src/matched_filter/ipp_matched_filter.hpp:60-90)
class HipMatchedFilter : public MatchedFilter<HipMatchedFilter> {
protected:
    friend class MatchedFilter<HipMatchedFilter>;

    template <typename T>
    static std::vector<std::complex<float>> Prepare(const std::vector<T>& impulse_response) {
        std::vector<float> code(impulse_response.begin(), impulse_response.end());
        std::vector<std::complex<float>> spec(code.size());
        ugsdr_hip::check(ugsdr_hip_mf_prepare(hip_detail::ctx(), code.data(), hip_detail::cf(spec.data()),
                                              (int64_t)code.size()));
        return spec;
    }
    static void ProcessOptimized(std::vector<std::complex<float>>& src_dst,
                                 const std::vector<std::complex<float>>& spectrum) {
        if (spectrum.size() != src_dst.size()) throw std::runtime_error("Size mismatch");
        ugsdr_hip::check(ugsdr_hip_mf_filter(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                             hip_detail::cf(spectrum.data()), hip_detail::cf(src_dst.data()),
                                             (int64_t)src_dst.size()));
    }
    template <typename T>
    static void Process(std::vector<std::complex<float>>& src_dst, const std::vector<T>& impulse_response) {
        ProcessOptimized(src_dst, Prepare(impulse_response));
    }
    template <typename U, typename S>
        requires(!hip_detail::is_fp32<U>)
    static void ProcessOptimized(std::vector<std::complex<U>>&, const S&) {
        hip_detail::fp32_only<U>();
    }
    template <typename U, typename T>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<std::complex<U>>&, const std::vector<T>&) {
        hip_detail::fp32_only<U>();
    }
};

// IppAbs counterpart (src/math/ipp_abs.hpp:25-33)
class HipAbs : public Abs<HipAbs> {
protected:
    friend class Abs<HipAbs>;
    static std::vector<float> Process(const std::vector<std::complex<float>>& src) {
        std::vector<float> dst(src.size());
        ugsdr_hip::check(ugsdr_hip_abs(hip_detail::ctx(), hip_detail::cf(src.data()), dst.data(), (int64_t)src.size()));
        return dst;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static std::vector<U> Process(const std::vector<std::complex<U>>& src) {
        hip_detail::fp32_only<U>();
        return std::vector<U>(src.size());
    }
};

// IppReshapeAndSum counterpart (src/math/ipp_reshape_and_sum.hpp:23-53)
class HipReshapeAndSum : public ReshapeAndSum<HipReshapeAndSum> {
protected:
    friend class ReshapeAndSum<HipReshapeAndSum>;
    static void Process(std::vector<float>& src_dst, std::size_t block) {
        std::vector<float> out(block);
        ugsdr_hip::check(ugsdr_hip_reshape_sum(hip_detail::ctx(), src_dst.data(), out.data(),
                                               (int64_t)src_dst.size(), (int64_t)block, 0));
        src_dst = std::move(out);
    }
    static void Process(std::vector<std::complex<float>>& src_dst, std::size_t block) {
        std::vector<std::complex<float>> out(block);
        ugsdr_hip::check(ugsdr_hip_reshape_sum(hip_detail::ctx(), hip_detail::cf(src_dst.data()),
                                               hip_detail::cf(out.data()), (int64_t)src_dst.size(), (int64_t)block,
                                               1));
        src_dst = std::move(out);
    }
    template <typename T>
    static auto Process(const std::vector<T>& src, std::size_t block) {
        auto dst = src;
        Process(dst, block);
        return dst;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<U>&, std::size_t) {
        hip_detail::fp32_only<U>();
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<std::complex<U>>&, std::size_t) {
        hip_detail::fp32_only<U>();
    }
};

// IppMaxIndex counterpart (src/math/ipp_max_index.hpp:27-39): first index of the max
class HipMaxIndex : public MaxIndex<HipMaxIndex> {
protected:
    friend class MaxIndex<HipMaxIndex>;
    static MaxIndex::Result<float> Process(const std::vector<float>& src) {
        float v = 0.f;
        int64_t i = 0;
        ugsdr_hip::check(ugsdr_hip_max_index(hip_detail::ctx(), src.data(), (int64_t)src.size(), &v, &i));
        MaxIndex::Result<float> r;
        r.value = v;
        r.index = (std::size_t)i;
        return r;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static MaxIndex::Result<U> Process(const std::vector<U>&) {
        hip_detail::fp32_only<U>();
        return {};
    }
};

// IppMeanStdDev counterpart (src/math/ipp_mean_stddev.hpp:27-39): sample sigma
class HipMeanStdDev : public MeanStdDev<HipMeanStdDev> {
protected:
    friend class MeanStdDev<HipMeanStdDev>;
    static MeanStdDev::Result<float> Process(const std::vector<float>& src) {
        MeanStdDev::Result<float> r;
        ugsdr_hip::check(ugsdr_hip_mean_std(hip_detail::ctx(), src.data(), (int64_t)src.size(), &r.mean, &r.sigma));
        return r;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static MeanStdDev::Result<U> Process(const std::vector<U>&) {
        hip_detail::fp32_only<U>();
        return {};
    }
};

// IppResampler counterpart (src/resample/ipp_resampler.hpp:131-143)
class HipResampler : public Resampler<HipResampler> {
protected:
    friend class Resampler<HipResampler>;
    static void Process(std::vector<std::complex<float>>& src_dst, std::size_t new_fs, std::size_t old_fs) {
        const int64_t n = ugsdr_resample_len((int64_t)src_dst.size(), (int64_t)new_fs, (int64_t)old_fs);
        if (n < 0) throw std::runtime_error("bad sampling rates");
        std::vector<std::complex<float>> out((std::size_t)n);
        ugsdr_hip::check(ugsdr_hip_resample(hip_detail::ctx(), hip_detail::cf(src_dst.data()), (int64_t)src_dst.size(),
                                            (int64_t)new_fs, (int64_t)old_fs, hip_detail::cf(out.data())));
        src_dst = std::move(out);
    }
    static auto Process(const std::vector<std::complex<float>>& src, std::size_t new_fs, std::size_t old_fs) {
        auto dst = src;
        Process(dst, new_fs, old_fs);
        return dst;
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static void Process(std::vector<std::complex<U>>&, std::size_t, std::size_t) {
        hip_detail::fp32_only<U>();
    }
    template <typename U>
        requires(!hip_detail::is_fp32<U>)
    static std::vector<std::complex<U>> Process(const std::vector<std::complex<U>>& src, std::size_t, std::size_t) {
        hip_detail::fp32_only<U>();
        return src;
    }
};

// SequentialUpsampler counterpart (src/resample/upsampler.hpp:31-37)
class HipUpsampler : public Upsampler<HipUpsampler> {
protected:
    friend class Upsampler<HipUpsampler>;
    // codes of any integer or float element type; double values would be
    // narrowed on the way through the device, so they stop at compile time
    template <typename T>
    static void Process(std::vector<T>& src_dst, std::size_t samples) {
        if constexpr (std::is_floating_point_v<T>) hip_detail::fp32_only<T>();
        std::vector<float> in(src_dst.begin(), src_dst.end()), out(samples);
        ugsdr_hip::check(ugsdr_hip_upsample(hip_detail::ctx(), in.data(), (int64_t)in.size(), out.data(),
                                            (int64_t)samples));
        src_dst.assign(out.begin(), out.end());
    }
};

// IppCorrelator counterpart (src/correlator/ipp_correlator.hpp:25-38)
class HipCorrelator : public Correlator<HipCorrelator> {
protected:
    friend class Correlator<HipCorrelator>;
    template <typename T>
    static std::complex<float> Process(const std::span<const std::complex<float>>& signal,
                                       const std::span<const T>& code) {
        if (signal.size() != code.size()) throw std::runtime_error("Size mismatch");
        std::vector<float> c(code.begin(), code.end());
        std::complex<float> out{};
        ugsdr_hip::check(ugsdr_hip_correlate(hip_detail::ctx(), hip_detail::cf(signal.data()),
                                             (int64_t)signal.size(), c.data(), (int64_t)c.size(),
                                             hip_detail::cf(&out)));
        return out;
    }
    template <typename T1, typename T2>
    static auto Process(const T1& signal, const T2& code) {
        hip_detail::fp32_only<typename T1::value_type::value_type>();
        return Process(std::span<const std::complex<float>>(signal.data(), signal.size()),
                       std::span<const typename T2::value_type>(code.data(), code.size()));
    }
};

}  // namespace ugsdr
