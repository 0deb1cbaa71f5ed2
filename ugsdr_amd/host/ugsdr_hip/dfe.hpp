// Digital front-end channel on the GPU: the counterpart of ugsdr::Channel
// (src/dfe/dfe.hpp:110-203) for a recording held in memory.  The recording's
// bytes (any UGSDR_FMT_*, packed BbpDdc / Nt1065 included) are decoded, mixed
// from the file's centre frequency to the signal's with a phase that runs on
// across calls, and resampled, in one device pass per call.
#pragma once

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <vector>

#include "context.hpp"
#include "fse.hpp"
#include "jse.hpp"

namespace ugsdr {

// Channel::CentralFrequency (src/dfe/dfe.hpp:118-157): carrier of each signal
inline double CentralFrequency(Signal signal) {
    switch (signal) {
        case Signal::GpsCoarseAcquisition_L1:
        case Signal::Galileo_E1b:
        case Signal::Galileo_E1c:
        case Signal::BeiDou_B1C:
        case Signal::SbasCoarseAcquisition_L1:
        case Signal::QzssCoarseAcquisition_L1:
        case Signal::Qzss_L1S: return 1575.42e6;
        case Signal::GlonassCivilFdma_L1: return 1602e6;
        case Signal::GlonassCivilFdma_L2: return 1602e6 * 7 / 9;
        case Signal::BeiDou_B1I: return 1561.098e6;
        case Signal::Gps_L2CM:
        case Signal::Qzss_L2CM: return 1227.6e6;
        case Signal::Gps_L5I:
        case Signal::Gps_L5Q:
        case Signal::Galileo_E5aI:
        case Signal::Galileo_E5aQ:
        case Signal::NavIC_L5:
        case Signal::Sbas_L5I:
        case Signal::Sbas_L5Q:
        case Signal::Qzss_L5I:
        case Signal::Qzss_L5Q: return 1176.45e6;
        case Signal::Galileo_E5bI:
        case Signal::Galileo_E5bQ: return 1207.14e6;
        case Signal::Galileo_E6b:
        case Signal::Galileo_E6c: return 1278.75e6;
    }
    throw std::runtime_error("Unexpected signal");
}

class HipDfeChannel {
public:
    double sampling_rate = 0.0;      // the channel's output rate
    double central_frequency = 0.0;  // the signal's carrier
    bool spectrum_inversion = false;

    // Channel(signal_params, signal, new_sampling_rate) (dfe.hpp:163-174): the
    // mixer moves the file's centre frequency onto the signal's carrier.
    // mitigation: Config::IsMitigationEnabled() (dfe.hpp:200-201), the
    // interference suppression run on each call's output
    HipDfeChannel(int fmt, double file_sampling_rate, double file_central_frequency, Signal signal,
                  double new_sampling_rate, ugsdr_hip::HipContext& ctx = ugsdr_hip::HipContext::Default(),
                  bool mitigation = false)
        : sampling_rate(new_sampling_rate), central_frequency(CentralFrequency(signal)),
          spectrum_inversion(file_central_frequency - central_frequency > 1),
          output_per_ms((std::size_t)(new_sampling_rate / 1e3)), mitigate(mitigation),
          jamming_suppressor(new_sampling_rate, UGSDR_STFT_BLACKMAN_HARRIS, ctx) {
        ugsdr_hip::check(ugsdr_hip_dfe_create(ctx.get(), fmt, file_sampling_rate, new_sampling_rate,
                                              file_central_frequency - central_frequency, 0.0, &ch));
    }
    ~HipDfeChannel() { ugsdr_hip_dfe_destroy(ch); }
    HipDfeChannel(const HipDfeChannel&) = delete;
    HipDfeChannel& operator=(const HipDfeChannel&) = delete;

    // Channel::GetSeveralEpochs (dfe.hpp:192-203) without interference
    // suppression: epoch_cnt ms of the recording starting at `bytes`
    std::vector<std::complex<float>> GetSeveralEpochs(const void* bytes, std::size_t epoch_cnt) {
        std::vector<std::complex<float>> out(epoch_cnt * output_per_ms);
        ugsdr_hip::check(ugsdr_hip_dfe_process(ch, bytes, 0, (int64_t)epoch_cnt, reinterpret_cast<float*>(out.data()), 0));
        if (mitigate) jamming_suppressor.Process(out);
        return out;
    }
    // the same, written to device memory for a later device-side search
    void GetSeveralEpochsOnDevice(const void* bytes, std::size_t epoch_cnt, std::complex<float>* device_out) {
        ugsdr_hip::check(
            ugsdr_hip_dfe_process(ch, bytes, 0, (int64_t)epoch_cnt, reinterpret_cast<float*>(device_out), 1));
        if (mitigate) jamming_suppressor.ProcessOnDevice(device_out, epoch_cnt * output_per_ms);
    }
    const auto& GetDetectionResults() const { return jamming_suppressor.GetDetectionResults(); }
    std::vector<std::complex<float>> GetEpoch(const void* bytes) { return GetSeveralEpochs(bytes, 1); }

    double GetPhase() const {
        double p = 0;
        ugsdr_hip::check(ugsdr_hip_dfe_phase(ch, &p));
        return p;
    }

private:
    std::size_t output_per_ms = 0;
    bool mitigate = false;
    HipJammingSuppressionEngine jamming_suppressor;
    ugsdr_hip_dfe* ch = nullptr;
};

}  // namespace ugsdr
