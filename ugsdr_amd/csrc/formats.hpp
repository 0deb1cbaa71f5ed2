// Recorded sample formats: device decoding and host-side geometry.
//
// Mirrors SignalParametersBase::GetPartialSignal (src/signal_parameters.hpp:
// 140-237) for the FileType values of signal_parameters.hpp:23-32.  Decoding
// happens where the samples are first read (the acquisition front kernel, the
// unpack operator), so packed recordings travel to the device as they are
// stored.
//   CI8 / CI16 / CF32 : interleaved I/Q (Iq_8_plus_8, Iq_16_plus_16, float)
//   R8                : real int8 (Real_8), imaginary part 0
//   BBP_DDC           : epochs of one 8-byte header (a ms counter in its low
//                       24 bits) followed by 2+2-bit complex samples, two per
//                       byte: sample 2j in bits 0-3 (im 0-1, re 2-3), sample
//                       2j+1 in bits 4-7 (im 4-5, re 6-7).  Each 2-bit field is
//                       a signed value v and the sample component is (v | 1),
//                       i.e. -1 when the field's top bit is set, else +1
//                       (BbpPackedSpan.hpp:11-19).  The epoch length follows
//                       signal_parameters.hpp:74-80.
//   NT1065_1..4       : one byte per sample holding four channels; channel c
//                       uses bits 6-2c (sign) and 7-2c (magnitude): the real
//                       value (2 sign - 1)(1 + 2 magnitude) in {-3,-1,1,3}
//                       (NtlabPackedSpan.hpp:10-21, offsets :92-99).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ugsdr_hip.h"

namespace ugsdr_hip {

struct SampleSource {
    const void* p;
    int fmt;
    int spm;          // BBP_DDC: samples per epoch (1 ms)
    int epoch_bytes;  // BBP_DDC: bytes per epoch, header included
};

__device__ __forceinline__ float2 load_sample(const SampleSource& s, int64_t i) {
    switch (s.fmt) {
        case UGSDR_FMT_CI8: {
            const char2 v = static_cast<const char2*>(s.p)[i];
            return make_float2((float)v.x, (float)v.y);
        }
        case UGSDR_FMT_CI16: {
            const short2 v = static_cast<const short2*>(s.p)[i];
            return make_float2((float)v.x, (float)v.y);
        }
        case UGSDR_FMT_CF32: return static_cast<const float2*>(s.p)[i];
        case UGSDR_FMT_R8: return make_float2((float)static_cast<const signed char*>(s.p)[i], 0.f);
        case UGSDR_FMT_BBP_DDC: {
            // sample counts of a request stay below 2^31 (checked on the host)
            const uint32_t e = (uint32_t)i / (uint32_t)s.spm, j = (uint32_t)i - e * (uint32_t)s.spm;
            const uint32_t b = static_cast<const unsigned char*>(s.p)[(size_t)e * s.epoch_bytes + 8 + (j >> 1)] >>
                               ((j & 1u) * 4u);
            return make_float2((b & 8u) ? -1.f : 1.f, (b & 2u) ? -1.f : 1.f);
        }
        default: {  // NT1065_1..4: channel bit offsets 6, 4, 2, 0
            const uint32_t v = (uint32_t)static_cast<const unsigned char*>(s.p)[i] >> (2 * (UGSDR_FMT_NT1065_4 - s.fmt));
            const float m = (v & 2u) ? 3.f : 1.f;
            return make_float2((v & 1u) ? m : -m, 0.f);
        }
    }
}

// load_sample with the format fixed at compile time for the interleaved I/Q
// formats (straight-line loads the compiler can issue back to back); any
// other FMT value decodes through the run-time switch.
template <int FMT>
__device__ __forceinline__ float2 load_sample_fmt(const SampleSource& s, int64_t i) {
    if constexpr (FMT == UGSDR_FMT_CI8) {
        const char2 v = static_cast<const char2*>(s.p)[i];
        return make_float2((float)v.x, (float)v.y);
    } else if constexpr (FMT == UGSDR_FMT_CI16) {
        const short2 v = static_cast<const short2*>(s.p)[i];
        return make_float2((float)v.x, (float)v.y);
    } else if constexpr (FMT == UGSDR_FMT_CF32) {
        return static_cast<const float2*>(s.p)[i];
    } else {
        return load_sample(s, i);
    }
}

// The front end of every path that reads recorded samples: decode, mix by a
// tone (Mixer::Translate, src/mixer/mixer.hpp:33-51, FP64 phase), and the
// rational resampler (Resampler::Transform, src/resample/resampler.hpp:39-46:
// sample-and-hold up to lcm, boxcar down), one output sample per thread.
struct FrontArgs {
    SampleSource src;  // the samples in their recorded format
    int64_t n_in;
    int64_t n_out;
    double f_turns;   // mixing frequency / fs_in (cycles per sample), 0 = no tone
    double ph_turns;  // initial phase in cycles
    int apply_tone;
    int64_t up;       // lcm / fs_in
    int64_t down;     // lcm / fs_out
    float2* out;
};
int launch_front_args(const FrontArgs& fa, hipStream_t st);

// Geometry of a format at input rate fs: fills spm / epoch_bytes (BBP_DDC
// needs fs to be a whole number of kHz) and returns 0, or a UGSDR_ERR_* code
// with the error text set.
int format_geometry(int fmt, double fs, int& spm, int& epoch_bytes);
// bytes holding n samples (BBP_DDC: whole epochs, headers included)
int64_t format_bytes(int fmt, int64_t n, int spm, int epoch_bytes);
// SignalParametersBase::VerifyHeaders (signal_parameters.hpp:36-45): the ms
// counters of consecutive epochs must step by one.  0 or UGSDR_ERR_ARG.
int verify_bbp_headers(const void* host, int64_t n_epochs, int epoch_bytes);

}  // namespace ugsdr_hip
