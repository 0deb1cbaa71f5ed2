// Interference suppression on the GPU: the counterpart of
// ugsdr::JammingSuppressionEngine (src/antijamming/jse.hpp:25-110) with its
// JammingDetector (src/antijamming/jamming_detection.hpp:25-214).
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

#include "backends.hpp"
#include "context.hpp"

namespace ugsdr {

// IppStft / SequentialStft counterpart (src/math/ipp_stft.hpp:14-96,
// src/math/stft.hpp:31-88): the window is IppStft's Blackman-Harris or
// SequentialStft's Hann; the device transform has the reference's default
// geometry (128-sample windows, 96 overlap).
template <int Window = UGSDR_STFT_BLACKMAN_HARRIS>
class HipStftT : public ShortTimeFourierTransform<HipStftT<Window>> {
protected:
    friend class ShortTimeFourierTransform<HipStftT<Window>>;
    static void CheckGeometry(std::size_t window_length, std::size_t overlap_length) {
        if (window_length != 128 || overlap_length != 96)
            throw std::runtime_error("HipStft runs 128-sample windows with 96 overlap");
    }
    static auto Process(const std::vector<std::complex<float>>& src, std::size_t window_length,
                        std::size_t overlap_length) {
        CheckGeometry(window_length, overlap_length);
        const std::size_t ncol = src.size() >= 128 ? (src.size() - 96) / 32 : 0;
        std::vector<std::complex<float>> flat(ncol * 128);
        ugsdr_hip::check(ugsdr_hip_stft(hip_detail::ctx(), hip_detail::cf(src.data()), (int64_t)src.size(), Window,
                                        hip_detail::cf(flat.data())));
        std::vector<std::vector<std::complex<float>>> dst(ncol);
        for (std::size_t c = 0; c < ncol; ++c) dst[c].assign(flat.begin() + c * 128, flat.begin() + (c + 1) * 128);
        return dst;
    }
    static auto Process(const std::vector<std::vector<std::complex<float>>>& src, std::size_t window_length,
                        std::size_t overlap_length) {
        CheckGeometry(window_length, overlap_length);
        std::vector<std::complex<float>> flat;
        for (auto& col : src) {
            if (col.size() != 128) throw std::runtime_error("Size mismatch");
            flat.insert(flat.end(), col.begin(), col.end());
        }
        std::vector<std::complex<float>> dst(src.empty() ? 0 : 128 + (src.size() - 1) * 32);
        if (!src.empty())
            ugsdr_hip::check(ugsdr_hip_istft(hip_detail::ctx(), hip_detail::cf(flat.data()), (int64_t)src.size(),
                                             hip_detail::cf(dst.data())));
        return dst;
    }
};
using HipStft = HipStftT<UGSDR_STFT_BLACKMAN_HARRIS>;
using HipSequentialStft = HipStftT<UGSDR_STFT_HANN>;


class HipJammingSuppressionEngine {
public:
    // JammingDetector::DetectionResult (jamming_detection.hpp:27-48)
    struct DetectionResult {
        enum class InterferenceType : std::size_t { None = 0, Narrowband, Wideband };
        InterferenceType interference_type = InterferenceType::None;
        std::size_t frequency_bin = 0;
        double frequency = 0.0;
        double jammer_to_noise = 0.0;
        bool is_wideband() const { return interference_type == InterferenceType::Wideband; }
    };

    // window: UGSDR_STFT_BLACKMAN_HARRIS (the IPP build's IppStft) or
    // UGSDR_STFT_HANN (SequentialStft)
    explicit HipJammingSuppressionEngine(double sampling_rate, int window = UGSDR_STFT_BLACKMAN_HARRIS,
                                         ugsdr_hip::HipContext& ctx = ugsdr_hip::HipContext::Default())
        : fs(sampling_rate), window_kind(window), context(&ctx) {}

    // Process (jse.hpp:93-103): detection, then the notch or the excision, in place
    void Process(std::vector<std::complex<float>>& src_dst) {
        Run(reinterpret_cast<float*>(src_dst.data()), src_dst.size(), false);
    }
    // the same on n samples already in device memory
    void ProcessOnDevice(std::complex<float>* device_samples, std::size_t n) {
        Run(reinterpret_cast<float*>(device_samples), n, true);
    }

    const auto& GetDetectionResults() const { return detection_results; }

private:
    double fs;
    int window_kind;
    ugsdr_hip::HipContext* context;
    std::vector<DetectionResult> detection_results;

    void Run(float* samples, std::size_t n, bool on_device) {
        ugsdr_jammer found[32];
        int32_t count = 0;
        ugsdr_hip::check(ugsdr_hip_jse_process(context->get(), samples, (int64_t)n, fs, window_kind, on_device ? 1 : 0,
                                               found, 32, &count));
        detection_results.clear();
        for (int32_t i = 0; i < count && i < 32; ++i) {
            DetectionResult d;
            d.interference_type = found[i].type == UGSDR_JAMMER_WIDEBAND ? DetectionResult::InterferenceType::Wideband
                                                                         : DetectionResult::InterferenceType::Narrowband;
            d.frequency_bin = (std::size_t)found[i].frequency_bin;
            d.frequency = found[i].frequency;
            d.jammer_to_noise = found[i].jammer_to_noise;
            detection_results.push_back(d);
        }
    }
};

}  // namespace ugsdr
