// Interference suppression on the GPU: the counterpart of
// ugsdr::JammingSuppressionEngine (src/antijamming/jse.hpp:25-110) with its
// JammingDetector (src/antijamming/jamming_detection.hpp:25-214).
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

#include "context.hpp"

namespace ugsdr {

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
