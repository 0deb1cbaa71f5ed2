// C++20 host surface test: the reference's unit KATs through the Hip* CRTP
// backends, the reference's real_8 acquisition KAT through both the
// per-operator FastSearchEngine<HipFseConfig> loop and the fused
// HipFastSearchEngine, and the fused engine straight from the packed BbpDdc
// recording.  argv[1]: tests/golden/real_8_first4ms.bin; argv[2]:
// tests/golden/bbp_ddc_gps_L1_first4ms.bin.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <complex>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <fstream>
#include <map>
#include <utility>
#include <numeric>
#include <vector>

#include "ugsdr_hip/dfe.hpp"
#include "ugsdr_hip/fse.hpp"
#include "ugsdr_hip/tracking.hpp"
#include "ugsdr_hip/jse.hpp"

#define EXPECT(cond)                                                                \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            return 1;                                                               \
        }                                                                           \
    } while (0)

using cvec = std::vector<std::complex<float>>;

static std::vector<int8_t> l1ca(int sv) {
    std::vector<int8_t> c(1023);
    ugsdr_hip::check(ugsdr_code_gps_l1ca(sv, c.data()));
    return c;
}

static int unit_kats() {
    // mixer, test/basic_tests.cpp:321-350
    ugsdr::HipMixer mixer(1e6, 1e6 / 2, 0);
    cvec ones(1000, 1.0f);
    mixer.Translate(ones);
    for (std::size_t i = 0; i < ones.size(); ++i) {
        EXPECT(ones[i].real() == 1.0f - 2.0f * (float)(i & 1));
        EXPECT(std::abs(ones[i].imag()) < 1e-6f);
    }
    // matched filter, :368-379 (GLONASS m-sequence)
    std::vector<int8_t> g(511);
    ugsdr_hip::check(ugsdr_code_glonass_of(g.data()));
    std::vector<float> code(g.begin(), g.end());
    const cvec sig(code.begin(), code.end());
    auto mf = ugsdr::HipMatchedFilter::Filter(sig, code);
    EXPECT(std::abs(mf[0].real() - 511.0f) < 5e-3f);
    for (std::size_t i = 1; i < mf.size(); ++i) EXPECT(std::abs(mf[i].real() + 1.0f) < 5e-3f);
    // correlator self-dot, :72-86
    auto c0 = l1ca(0);
    std::vector<float> cf(c0.begin(), c0.end());
    cvec cs(cf.begin(), cf.end());
    auto dot = ugsdr::HipCorrelator::Correlate(std::span(cs), std::span(cf));
    EXPECT(dot == std::complex<float>(1023.0f, 0.0f));
    bool threw = false;
    try {
        (void)ugsdr::HipCorrelator::Correlate(std::span(cs).subspan(0, 10), std::span(cf));
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    // abs, max index, mean/std, reshape, :408-433, :528-553, :607-630
    auto a = ugsdr::HipAbs::Transform(cvec(1000, {1.0f, -1.0f}));
    for (float v : a) EXPECT(std::abs(v - std::sqrt(2.0f)) < 1e-2f);
    std::vector<float> iota(1000);
    std::iota(iota.begin(), iota.end(), 0.0f);
    auto mx = ugsdr::HipMaxIndex::Transform(iota);
    EXPECT(mx.value == 999.0f && mx.index == 999);
    auto ms = ugsdr::HipMeanStdDev::Calculate(std::vector<float>{1, 2, 3, 4});
    EXPECT(std::abs(ms.mean - 2.5f) < 1e-6f && std::abs(ms.sigma - 1.2909944f) < 1e-5f);
    auto rs = ugsdr::HipReshapeAndSum::Transform(std::vector<float>(10000, 1.0f), 1000);
    EXPECT(rs.size() == 1000 && rs[17] == 10.0f);
    threw = false;
    try {
        (void)ugsdr::HipReshapeAndSum::Transform(std::vector<float>(10, 1.0f), 3);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    // upsampler x10, :800-817
    auto up = ugsdr::HipUpsampler::Transform(std::as_const(iota), 10000);
    EXPECT(up.size() == 10000 && up[12] == 1.0f && up[9999] == 999.0f);
    std::puts("unit KATs ok");
    return 0;
}

static int real8_kat(const char* path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<int8_t> raw(4 * 38192);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    EXPECT(f.gcount() == (std::streamsize)raw.size());
    const double fs = 38.192e6;
    // DFE channel: real samples, member Translate by +9.55 MHz (src/dfe/dfe.hpp:199)
    cvec x(raw.begin(), raw.end());
    ugsdr::HipMixer dfe_mixer(fs, 9.55e6, 0);
    dfe_mixer.Translate(x);
    std::vector<ugsdr::Sv> svs;
    std::vector<std::vector<float>> codes_f;
    std::vector<std::vector<int8_t>> codes;
    for (int s = 0; s < 32; ++s) {
        svs.push_back({s, ugsdr::System::Gps, ugsdr::Signal::GpsCoarseAcquisition_L1});
        codes.push_back(l1ca(s));
        codes_f.emplace_back(codes.back().begin(), codes.back().end());
    }
    ugsdr::SearchEpoch epoch{x.data(), x.size(), fs, 0.0};
    ugsdr::FastSearchEngine<ugsdr::HipFseConfig> per_op(6e3, 200);
    auto r1 = per_op.Process(epoch, svs, codes_f, 4);
    // fused: raw int8 with the DFE tone as the request's mix
    ugsdr::HipFastSearchEngine::Options opt;
    ugsdr::HipFastSearchEngine fused(6e3, 200, opt);
    auto r2 = fused.Process(x.data(), UGSDR_FMT_CF32, x.size(), fs, 0.0, svs, codes);
    std::printf("per-operator %zu detections, fused %zu detections\n", r1.size(), r2.size());
    EXPECT(r1.size() == r2.size());
    for (std::size_t i = 0; i < r1.size(); ++i) {
        EXPECT(r1[i].sv_number.id == r2[i].sv_number.id);
        EXPECT(r1[i].doppler == r2[i].doppler);
        EXPECT(std::abs(r1[i].code_offset - r2[i].code_offset) <= 4.0);
        EXPECT(std::abs(r1[i].GetSnr() - r2[i].GetSnr()) <= 1e-3 * r1[i].GetSnr());
        EXPECT(r2[i].output_peak.size() == 9548);
    }
    // test/integration_tests.cpp:253-263 (SV ids 0-based), >= 4 within 400 Hz (:314)
    const std::map<int, double> kat = {{2, 0}, {5, 6000}, {8, -1000}, {14, 0}, {17, 2000}, {20, 3000}, {21, 0}, {25, 5000}};
    int match = 0;
    for (auto& r : r2) {
        auto it = kat.find(r.sv_number.id);
        if (it != kat.end() && std::abs(r.doppler - it->second) <= 400) ++match;
    }
    std::printf("real_8 KAT: %d/8 matched\n", match);
    EXPECT(match >= 4);
    return 0;
}

static int bbp_recording(const char* path) {
    const double fs = 33.25e6;
    const int64_t spm = 33250;
    const int64_t eb = ugsdr_format_bytes(UGSDR_FMT_BBP_DDC, spm, fs);
    EXPECT(eb == 16640);
    std::ifstream f(path, std::ios::binary);
    std::vector<uint8_t> raw(4 * (std::size_t)eb);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    EXPECT(f.gcount() == (std::streamsize)raw.size());
    auto x = ugsdr_hip::GetPartialSignal(raw.data(), UGSDR_FMT_BBP_DDC, 4 * spm, fs);
    // the first two samples of each epoch: sign bits 3/1 and 7/5 of the byte
    // after the 8-byte header (src/helpers/BbpPackedSpan.hpp:11-19)
    for (int e = 0; e < 4; ++e) {
        const uint8_t b = raw[(std::size_t)e * eb + 8];
        EXPECT(x[e * spm] == std::complex<float>((b & 0x08) ? -1.f : 1.f, (b & 0x02) ? -1.f : 1.f));
        EXPECT(x[e * spm + 1] == std::complex<float>((b & 0x80) ? -1.f : 1.f, (b & 0x20) ? -1.f : 1.f));
    }
    auto bad = raw;
    bad[2 * (std::size_t)eb] ^= 1;
    bool threw = false;
    try {
        (void)ugsdr_hip::GetPartialSignal(bad.data(), UGSDR_FMT_BBP_DDC, 4 * spm, fs);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    // DFE channel: the recording is centred a little off L1 here (+2.5 kHz), so
    // the mixer runs; four 1 ms epochs continue the phase like one 4 ms call
    ugsdr::HipDfeChannel by_ms(UGSDR_FMT_BBP_DDC, fs, 1575.42e6 + 2.5e3, ugsdr::Signal::GpsCoarseAcquisition_L1, 4.2e6);
    ugsdr::HipDfeChannel whole(UGSDR_FMT_BBP_DDC, fs, 1575.42e6 + 2.5e3, ugsdr::Signal::GpsCoarseAcquisition_L1, 4.2e6);
    EXPECT(by_ms.central_frequency == 1575.42e6 && by_ms.spectrum_inversion);
    EXPECT(ugsdr::CentralFrequency(ugsdr::Signal::Gps_L5I) == 1176.45e6);
    cvec parts;
    for (int e = 0; e < 4; ++e) {
        auto p = by_ms.GetEpoch(raw.data() + (std::size_t)e * eb);
        EXPECT(p.size() == 4200);
        parts.insert(parts.end(), p.begin(), p.end());
    }
    auto all = whole.GetSeveralEpochs(raw.data(), 4);
    float peak = 0.f, diff = 0.f;
    for (std::size_t i = 0; i < all.size(); ++i) {
        peak = std::max(peak, std::abs(all[i]));
        diff = std::max(diff, std::abs(all[i] - parts[i]));
    }
    EXPECT(all.size() == parts.size() && diff <= 1e-5f * peak);
    // fused engine on the packed bytes with the default FseConfig rate rule
    // (33.25 MHz -> 6.65 MHz, 4 ms coherent: N = 26600 = 2^3 5^2 7 19)
    std::vector<ugsdr::Sv> svs;
    std::vector<std::vector<int8_t>> codes;
    for (int s = 0; s < 32; ++s) {
        svs.push_back({s, ugsdr::System::Gps, ugsdr::Signal::GpsCoarseAcquisition_L1});
        codes.push_back(l1ca(s));
    }
    ugsdr::HipFastSearchEngine engine(6e3, 125);
    auto found = engine.Process(raw.data(), UGSDR_FMT_BBP_DDC, 4 * spm, fs, 0.0, svs, codes);
    const std::map<int, double> strong = {{8, -500.0}, {22, 1500.0}};
    int match = 0;
    for (auto& r : found) {
        std::printf("bbp: sv %d doppler %.0f snr %.2f\n", r.sv_number.id, r.doppler, r.GetSnr());
        auto it = strong.find(r.sv_number.id);
        if (it != strong.end() && std::abs(r.doppler - it->second) <= 250) ++match;
    }
    EXPECT(match == 2);
    std::puts("bbp recording ok");
    return 0;
}

// HipTracker: a clean GPS L1 C/A signal (sv 4, +1500.3 Hz, delay 1234
// samples at 4.092 MHz) seeded 40 Hz and 0.6 samples off; after 1 s the
// carrier is within 3 Hz and the prompt's energy is in I.
static int tracker_lock() {
    const double fs = 4.092e6, f0 = 1500.3, tau = 1234.0;
    const int spm = 4092, n_ms = 1000;
    const auto code = l1ca(4);
    std::vector<float> iq(2 * (std::size_t)spm * n_ms);
    for (int k = 0; k < spm * n_ms; ++k) {
        const long chip = (long)std::floor((k - tau) * 1.023e6 / fs);
        const double c = code[(std::size_t)(((chip % 1023) + 1023) % 1023)];
        const double ph = 2 * M_PI * f0 * k / fs + 0.3;
        iq[2 * (std::size_t)k] = (float)(c * std::cos(ph));
        iq[2 * (std::size_t)k + 1] = (float)(c * std::sin(ph));
    }
    ugsdr::AcquisitionResult<float> acq;
    acq.sv_number = {4, ugsdr::System::Gps, ugsdr::Signal::GpsCoarseAcquisition_L1};
    acq.doppler = f0 + 40.0;
    acq.code_offset = tau + 0.6;
    ugsdr::HipTracker tracker(fs, {acq}, {code});
    tracker.Track(iq.data(), UGSDR_FMT_CF32, 600);
    tracker.Track(iq.data() + 2 * (std::size_t)spm * 600, UGSDR_FMT_CF32, 400);   // the loops carry on across calls
    const auto& ch = tracker.GetTrackingParameters()[0];
    EXPECT(ch.prompt.size() == 1000 && ch.state.epochs == 1000);
    EXPECT(std::abs(ch.frequencies.back() - f0) < 3.0);
    double i_sum = 0, q_sum = 0;
    for (std::size_t b = 800; b < 1000; ++b) {
        i_sum += std::abs(ch.prompt[b].real());
        q_sum += std::abs(ch.prompt[b].imag());
    }
    EXPECT(q_sum < 0.15 * i_sum);
    std::printf("tracker: %.2f Hz after 1 s, code phase %.2f\n", ch.frequencies.back(), ch.code_phases.back());
    return 0;
}

// ProcessGalileo: two satellites whose milliseconds carry sign permutations
// 2 and 3 of the four, plus one absent code; 4 ms at 4.092 MHz, complex
// float.  Both are found at their planted code phase, the absent one is not.
static int galileo_hypotheses() {
    const double fs = 4.092e6;
    const int spm = 4092, n = 4 * spm;
    const int perms[4][4] = {{1, 1, 1, 1}, {1, 1, 1, -1}, {1, 1, -1, -1}, {1, -1, -1, -1}};
    std::mt19937 gen(42);
    std::vector<std::vector<int8_t>> codes(3);
    for (auto& c : codes) {
        c.resize(8184);
        for (std::size_t i = 0; i < 4092; ++i) {
            const int8_t v = (gen() & 1) ? 1 : -1;
            c[2 * i] = v;
            c[2 * i + 1] = (int8_t)-v;
        }
    }
    const int planted[2] = {2, 3}, delay[2] = {777, 12001};
    const double dop[2] = {1250.0, -2750.0};
    std::vector<std::complex<float>> x(n);
    std::normal_distribution<double> noise(0.0, 1.0);
    for (int k = 0; k < n; ++k) {
        std::complex<double> v(noise(gen), noise(gen));
        for (int s = 0; s < 2; ++s) {
            const int m = ((k - delay[s]) % n + n) % n;
            const double c = codes[s][(std::size_t)m * 8184 / n] * perms[planted[s]][m / spm];
            v += 0.1 * c * std::polar(1.0, 2 * M_PI * dop[s] * k / fs);
        }
        x[k] = std::complex<float>(v);
    }
    ugsdr::HipFseOptions opt;
    opt.coherent_ms = 4;
    opt.target_sampling_rate = fs;
    ugsdr::HipFastSearchEngine engine(5000, 250, opt);
    std::vector<ugsdr::Sv> svs(3);
    for (int s = 0; s < 3; ++s) svs[s] = {s, ugsdr::System::Galileo, ugsdr::Signal::Galileo_E1b};
    const auto found = engine.ProcessGalileo(x.data(), UGSDR_FMT_CF32, n, fs, 0.0, svs, codes);
    EXPECT(found.size() == 2);
    for (int s = 0; s < 2 && s < (int)found.size(); ++s) {
        EXPECT(found[s].sv_number.id == s);
        EXPECT(std::abs(found[s].code_offset - delay[s]) < 0.5);
        EXPECT(std::abs(found[s].doppler - dop[s]) < 1.0);
        std::printf("galileo: sv %d code offset %.0f doppler %.0f snr %.2f\n", s, found[s].code_offset,
                    found[s].doppler, found[s].GetSnr());
    }
    return 0;
}

// HipStft (the reference's STFT TestPeak, basic_tests.cpp:664-690) and
// HipJammingSuppressionEngine on a CW jammer 21 dB over the noise: found at
// its bin, and its power at its frequency cut by more than 30 dB.
static int antijamming() {
    const double fs = 50e6;
    std::vector<std::complex<float>> tone(50000);
    for (std::size_t i = 0; i < tone.size(); ++i)
        tone[i] = std::polar(1.0f, (float)(2 * M_PI * 1e6 * (double)i / fs));
    const auto cols = ugsdr::HipStft::Transform(tone);
    std::size_t near = 0;
    for (auto& col : cols) {
        std::size_t best = 0;
        for (std::size_t k = 1; k < col.size(); ++k)
            if (std::abs(col[k]) > std::abs(col[best])) best = k;
        near += std::abs(best * fs / 128 - 1e6) <= fs / 128 / 2;
    }
    EXPECT(cols.size() == (50000 - 96) / 32 && near >= 0.95 * cols.size());
    const auto back = ugsdr::HipStft::Transform(cols);
    EXPECT(back.size() == 128 + (cols.size() - 1) * 32);

    const double fs2 = 4.092e6;
    const int n = 4 * 4092;
    std::mt19937 gen(3);
    std::normal_distribution<double> noise(0.0, 0.5);
    std::vector<std::complex<float>> x(n);
    const double fj = 12 * fs2 / 128;
    for (int k = 0; k < n; ++k)
        x[k] = std::complex<float>(std::complex<double>(noise(gen), noise(gen)) + std::polar(6.0, 2 * M_PI * fj * k / fs2));
    auto power_at = [&](const std::vector<std::complex<float>>& v) {
        std::complex<double> acc = 0;
        for (int k = 0; k < n; ++k) acc += std::complex<double>(v[k]) * std::polar(1.0, -2 * M_PI * fj * k / fs2);
        return std::norm(acc / (double)n);
    };
    const double before = power_at(x);
    ugsdr::HipJammingSuppressionEngine jse(fs2);
    jse.Process(x);
    const auto& det = jse.GetDetectionResults();
    EXPECT(!det.empty() && det[0].frequency_bin == 12 && std::abs(det[0].frequency - fj) < 1e-6);
    const double after = power_at(x);
    EXPECT(10 * std::log10(before / after) > 30);
    std::printf("antijamming: %zu detections, jammer cut by %.1f dB\n", det.size(), 10 * std::log10(before / after));
    return 0;
}


static std::vector<uint8_t> read_file(const char* path, std::size_t n) {
    std::ifstream f(path, std::ios::binary);
    std::vector<uint8_t> raw(n);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
    if (f.gcount() != (std::streamsize)raw.size()) throw std::runtime_error(std::string("short read: ") + path);
    return raw;
}

// FastSearchEngineBase::ProcessGps (fse.hpp:327-335) through the fused engine:
// * with an L1 band that holds satellites, only L1 C/A results come back;
// * with only the L5 recording (66.5 MHz, 1176.45 MHz), L1 is skipped and
//   L5I finds satellites 2, 5, 8 and 25 at 16.625 MHz (AcquireL5, complex fold);
// * with a synthetic L5 band that holds only an L5Q signal, L5I finds nothing
//   and the L5Q fallback finds it at its planted code phase and Doppler.
static int process_gps(const char* l1_path, const char* l5_path) {
    std::vector<ugsdr::Sv> svs;
    for (int s = 0; s < 32; ++s) svs.push_back({s, ugsdr::System::Gps, ugsdr::Signal::GpsCoarseAcquisition_L1});
    ugsdr::HipFastSearchEngine engine(5e3, 200);
    using Band = ugsdr::HipFastSearchEngine::Subband;

    const auto l1_raw = read_file(l1_path, 4 * 16640);
    const auto l5_raw = read_file(l5_path, 4 * 33264);
    Band l1{l1_raw.data(), UGSDR_FMT_BBP_DDC, 4 * 33250, 33.25e6, 1575.42e6};
    Band l5{l5_raw.data(), UGSDR_FMT_BBP_DDC, 4 * 66500, 66.5e6, 1176.45e6};
    auto both = engine.ProcessGps(l1, l5, svs);
    EXPECT(!both.empty());
    for (auto& r : both) EXPECT(r.sv_number.signal == ugsdr::Signal::GpsCoarseAcquisition_L1);

    auto only_l5 = engine.ProcessGps(Band{}, l5, svs);
    std::vector<int> ids;
    for (auto& r : only_l5) {
        EXPECT(r.sv_number.signal == ugsdr::Signal::Gps_L5I);
        ids.push_back(r.sv_number.id);
        std::printf("gps L5I: sv %d doppler %.0f code offset %.0f snr %.2f\n", r.sv_number.id, r.doppler,
                    r.code_offset, r.GetSnr());
    }
    EXPECT((ids == std::vector<int>{2, 5, 8, 25}));

    // L5Q only: sv 6, 4001 samples (2000.5 chips) in, +1800 Hz, 20.46 MHz complex float, 4 ms
    const double fs = 20.46e6, dop = 1800.0;
    const int n = 4 * 20460, delay = 4001;
    std::vector<int8_t> q(10230);
    ugsdr_hip::check(ugsdr_code_gps_l5q(6, q.data()));
    std::mt19937 gen(9);
    std::normal_distribution<double> noise(0.0, 1.0);
    std::vector<std::complex<float>> x(n);
    for (int k = 0; k < n; ++k) {
        const int m = ((k - delay) % 20460 + 20460) % 20460;
        x[k] = std::complex<float>(std::complex<double>(noise(gen), noise(gen)) +
                                   0.08 * q[(std::size_t)m / 2] * std::polar(1.0, 2 * M_PI * dop * k / fs));
    }
    Band synth{x.data(), UGSDR_FMT_CF32, (std::size_t)n, fs, 1176.45e6};
    auto only_q = engine.ProcessGps(Band{}, synth, svs);
    EXPECT(only_q.size() == 1);
    if (!only_q.empty()) {
        EXPECT(only_q[0].sv_number.id == 6 && only_q[0].sv_number.signal == ugsdr::Signal::Gps_L5Q);
        EXPECT(only_q[0].code_offset == delay && only_q[0].doppler == dop);
        std::printf("gps L5Q fallback: sv %d doppler %.0f code offset %.0f snr %.2f\n", only_q[0].sv_number.id,
                    only_q[0].doppler, only_q[0].code_offset, only_q[0].GetSnr());
    }
    std::puts("process gps ok");
    return 0;
}

// HipTableMixer through the reference's Mixer surface: the in-place static
// overload and the member overload over two calls equal one call with the
// member's phase advance (TableMixer, src/mixer/table_mixer.hpp).
static int table_mixer() {
    const double fs = 4.092e6, f = 1250.0;
    std::vector<std::complex<float>> a(8184, {1.0f, 0.0f});
    ugsdr::HipTableMixer::Translate(a, fs, f, 0.0);
    // fs/4 exactly: cos/sin at quarter turns
    std::vector<std::complex<float>> q(8, {1.0f, 0.0f});
    ugsdr::HipTableMixer::Translate(q, 4.0, 1.0, 0.0);
    EXPECT(std::abs(q[1] - std::complex<float>(0.f, 1.f)) < 1e-6f && std::abs(q[2] + 1.0f) < 1e-6f);
    ugsdr::HipTableMixer m(fs, f, 0.0);
    std::vector<std::complex<float>> b1(4092, {1.0f, 0.0f}), b2(4092, {1.0f, 0.0f});
    m.Translate(b1);
    m.Translate(b2);
    EXPECT(b1[0] == a[0] && b1[4091] == a[4091]);
    EXPECT(std::abs(b2[0] - std::polar(1.0f, (float)(2 * M_PI * 0.25))) < 0.1f);   // phase 2 pi * frac(1.25)
    std::puts("table mixer ok");
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s real_8_first4ms.bin bbp_ddc_gps_L1_first4ms.bin bbp_ddc_gps_L5_first4ms.bin\n",
                     argv[0]);
        return 2;
    }
    try {
        if (int rc = unit_kats()) return rc;
        if (int rc = real8_kat(argv[1])) return rc;
        if (int rc = bbp_recording(argv[2])) return rc;
        if (int rc = tracker_lock()) return rc;
        if (int rc = galileo_hypotheses()) return rc;
        if (int rc = antijamming()) return rc;
        if (int rc = table_mixer()) return rc;
        if (int rc = process_gps(argv[2], argv[3])) return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 1;
    }
    std::puts("host API ok");
    return 0;
}
