// C entry points over the reference's own header-only operators, compiled
// straight from the reference checkout (oracle/Makefile target `ref`,
// output oracle/_ref/libugsdr_ref.so).  TEST INFRASTRUCTURE ONLY: it lets the
// tests compare the oracle restatement (refcpu.c) and the product's host
// tables with the reference code itself, wherever that code builds here with
// g++ alone.  Nothing here is restated: each function only marshals plain
// arrays into the reference's std::vector / std::span calls.
//
// Reference units used (paths relative to the reference's src/):
//   prn_codes/GpsL1Ca.hpp, GpsL5I.hpp, GpsL5Q.hpp, GlonassOf.hpp (Codegen::Get)
//   mixer/table_mixer.hpp + mixer/nco.hpp (TableMixer, in-place and member Translate)
//   mixer/mixer.hpp (SequentialMixer)
//   math/abs.hpp, max_index.hpp, mean_stddev.hpp, reshape_and_sum.hpp, dft.hpp (no FFTW)
//   matched_filter/matched_filter.hpp (SequentialMatchedFilter)
//   correlator/correlator.hpp, resample/upsampler.hpp, resample/resampler.hpp
//
// <stdexcept> is included first because the reference headers throw
// std::runtime_error without including it themselves.
#include <stdexcept>

#include "prn_codes/GpsL1Ca.hpp"
#include "prn_codes/GpsL5I.hpp"
#include "prn_codes/GpsL5Q.hpp"
#include "prn_codes/GlonassOf.hpp"
#include "mixer/mixer.hpp"
#include "mixer/table_mixer.hpp"
#include "math/abs.hpp"
#include "math/dft.hpp"
#include "math/max_index.hpp"
#include "math/mean_stddev.hpp"
#include "math/reshape_and_sum.hpp"
#include "matched_filter/matched_filter.hpp"
#include "correlator/correlator.hpp"
#include "resample/upsampler.hpp"
#include "resample/resampler.hpp"

#include <complex>
#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

namespace {
using cf = std::complex<float>;

std::vector<cf> to_cvec(const float* in, int64_t n) {
    std::vector<cf> v(static_cast<size_t>(n));
    std::memcpy(v.data(), in, sizeof(cf) * static_cast<size_t>(n));
    return v;
}

void from_cvec(const std::vector<cf>& v, float* out) { std::memcpy(out, v.data(), sizeof(cf) * v.size()); }

template <typename Gen>
int code(int sv, int8_t* out) {
    try {
        Gen::Get(out, static_cast<std::size_t>(sv));
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}
}  // namespace

extern "C" {

int ref_gps_l1ca(int sv, int8_t* out1023) { return code<ugsdr::GpsL1Ca>(sv, out1023); }
int ref_gps_l5i(int sv, int8_t* out10230) { return code<ugsdr::GpsL5I>(sv, out10230); }
int ref_gps_l5q(int sv, int8_t* out10230) { return code<ugsdr::GpsL5Q>(sv, out10230); }
int ref_glonass_of(int8_t* out511) { return code<ugsdr::GlonassOf>(0, out511); }

// Mixer<TableMixer>::Translate, in-place static overload (mixer.hpp:33-41).
void ref_table_mix(float* io_c32, int64_t n, double fs, double f, double phase) {
    auto v = to_cvec(io_c32, n);
    ugsdr::TableMixer::Translate(v, fs, f, phase);
    from_cvec(v, io_c32);
}

// Mixer<TableMixer> member overload (mixer.hpp:53-60), n_calls consecutive
// calls of n samples each on one mixer object; io holds n_calls * n samples.
void ref_table_mix_member(float* io_c32, int64_t n, int64_t n_calls, double fs, double f, double phase) {
    ugsdr::TableMixer m(fs, f, phase);
    for (int64_t c = 0; c < n_calls; ++c) {
        auto v = to_cvec(io_c32 + 2 * c * n, n);
        m.Translate(v);
        from_cvec(v, io_c32 + 2 * c * n);
    }
}

// Mixer<SequentialMixer>::Translate, const overload (mixer.hpp:43-51).
void ref_sequential_mix(const float* in_c32, float* out_c32, int64_t n, double fs, double f, double phase) {
    const auto v = to_cvec(in_c32, n);
    auto r = ugsdr::SequentialMixer::Translate(v, fs, f, phase);
    from_cvec(r, out_c32);
}

void ref_abs(const float* in_c32, int64_t n, float* out) {
    const auto v = to_cvec(in_c32, n);
    auto r = ugsdr::SequentialAbs::Transform(v);
    std::memcpy(out, r.data(), sizeof(float) * r.size());
}

int ref_reshape_sum(const float* in, int64_t n, int64_t block, float* out) {
    try {
        const std::vector<float> v(in, in + n);
        auto r = ugsdr::SequentialReshapeAndSum::Transform(v, static_cast<std::size_t>(block));
        std::memcpy(out, r.data(), sizeof(float) * r.size());
        return static_cast<int>(r.size());
    } catch (const std::exception&) {
        return -1;
    }
}

void ref_max_index(const float* in, int64_t n, float* value, int64_t* index) {
    const std::vector<float> v(in, in + n);
    auto r = ugsdr::SequentialMaxIndex::Transform(v);
    *value = r.value;
    *index = static_cast<int64_t>(r.index);
}

void ref_mean_std(const float* in, int64_t n, float* mean, float* sigma) {
    const std::vector<float> v(in, in + n);
    auto r = ugsdr::SequentialMeanStdDev::Calculate(v);
    *mean = r.mean;
    *sigma = r.sigma;
}

// SequentialDft without FFTW: the naive O(N^2) transform (dft.hpp:100-112).
void ref_dft(const float* in_c32, int64_t n, int inverse, float* out_c32) {
    const auto v = to_cvec(in_c32, n);
    auto r = ugsdr::SequentialDft::Transform(v, inverse != 0);
    from_cvec(r, out_c32);
}

// MatchedFilter<SequentialMatchedFilter>::FilterOptimized (matched_filter.hpp:
// 32-42, 53-58) against a spectrum made by the oracle or the device.
void ref_mf_filter_optimized(const float* in_c32, const float* spec_c32, int64_t n, float* out_c32) {
    const auto x = to_cvec(in_c32, n);
    const auto s = to_cvec(spec_c32, n);
    auto r = ugsdr::SequentialMatchedFilter::FilterOptimized(x, s);
    from_cvec(r, out_c32);
}

int ref_correlate(const float* sig_c32, int64_t n_sig, const float* code, int64_t n_code, float* out_c32) {
    try {
        std::span<const cf> s(reinterpret_cast<const cf*>(sig_c32), static_cast<size_t>(n_sig));
        std::span<const float> c(code, static_cast<size_t>(n_code));
        const cf r = ugsdr::SequentialCorrelator::Correlate(s, c);
        out_c32[0] = r.real();
        out_c32[1] = r.imag();
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

void ref_upsample(const float* in, int64_t n, float* out, int64_t samples) {
    const std::vector<float> v(in, in + n);
    auto r = ugsdr::SequentialUpsampler::Transform(v, static_cast<std::size_t>(samples));
    std::memcpy(out, r.data(), sizeof(float) * r.size());
}

// Resampler<SequentialResampler>::Transform (resampler.hpp:39-46); returns the
// output length, out must hold at least that many complex samples.
int64_t ref_resample(const float* in_c32, int64_t n, int64_t new_fs, int64_t old_fs, float* out_c32, int64_t cap) {
    const auto v = to_cvec(in_c32, n);
    auto r = ugsdr::SequentialResampler::Transform(v, static_cast<std::size_t>(new_fs), static_cast<std::size_t>(old_fs));
    if (static_cast<int64_t>(r.size()) <= cap)
        from_cvec(r, out_c32);
    return static_cast<int64_t>(r.size());
}

}  // extern "C"
