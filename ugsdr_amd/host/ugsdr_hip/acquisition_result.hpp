// Acquisition result and satellite id types with the reference's fields
// (src/common.hpp:147-154, src/acquisition/acquisition_result.hpp:12-31),
// shared by the acquisition and tracking surfaces.  Inside a reference build
// (UGSDR_HIP_REFERENCE_BASES) the reference's own acquisition_result.hpp
// supplies them.
#pragma once

#include <compare>
#include <cstdint>
#include <vector>

namespace ugsdr {

#ifndef UGSDR_HIP_REFERENCE_BASES
enum class System : std::uint32_t { Gps = 0, Glonass, Galileo, BeiDou, NavIC, Sbas, Qzss };
enum class Signal : std::uint32_t {
    GpsCoarseAcquisition_L1, Gps_L2CM, Gps_L5I, Gps_L5Q, GlonassCivilFdma_L1, GlonassCivilFdma_L2, Galileo_E1b,
    Galileo_E1c, Galileo_E5aI, Galileo_E5aQ, Galileo_E5bI, Galileo_E5bQ, Galileo_E6b, Galileo_E6c, BeiDou_B1I,
    BeiDou_B1C, NavIC_L5, SbasCoarseAcquisition_L1, Sbas_L5I, Sbas_L5Q, QzssCoarseAcquisition_L1, Qzss_L1S,
    Qzss_L2CM, Qzss_L5I, Qzss_L5Q,
};
// satellite id (src/common.hpp:147-154)
struct Sv {
    std::int32_t id = 0;
    System system = System::Gps;
    Signal signal = Signal::GpsCoarseAcquisition_L1;
};

// src/acquisition/acquisition_result.hpp:12-31
template <typename T>
struct AcquisitionResult final {
    Sv sv_number;
    double doppler = 0;
    double code_offset = 0;
    double level = 0;
    double sigma = 1.0;
    double intermediate_frequency = 0;
    std::vector<T> output_peak;
    auto operator<=>(const AcquisitionResult& rhs) const { return level <=> rhs.level; }
    auto GetSnr() const { return level / sigma; }
    auto GetAcquiredSignalType() const { return sv_number.signal; }
};

#endif

}  // namespace ugsdr
