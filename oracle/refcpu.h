/*
 * refcpu — CPU restatement of ugsdr's acquisition hot path and tracking
 * correlator taps.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product path (ugsdr_amd/) never links or calls it.
 *
 * Every function states the reference file:line it restates.  Paths are
 * relative to the ugsdr checkout.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   codes      : ICD first-10-chip octal words, Gold cross-correlation set
 *                {-65,-1,63}, GLONASS m-sequence autocorrelation 511/-1
 *                (test/basic_tests.cpp:368-379), self-dot 1023
 *                (test/basic_tests.cpp:72-86), L5 XB-advance table spot checks
 *   fft        : numpy.fft (pocketfft, FP64) at every factorisation used, the
 *                tone KAT (test/basic_tests.cpp:482-517), brute-force DFT
 *   mixer      : fs/2 KAT (test/basic_tests.cpp:321-350), closed form
 *   resample   : slope/boxcar KAT and x10 sample-and-hold KAT
 *                (test/basic_tests.cpp:764-850), integer-index formula
 *   abs/fold/max/meanstd : KATs test/basic_tests.cpp:408-433,528-553,607-630
 *                and numpy closed forms
 *   acquire    : planted truth of the seeded generator, numpy FP64
 *                restatement on small cases, the real_8 recording KAT
 *                (test/integration_tests.cpp:253-263)
 *   correlate/split : self-dot KAT, numpy dot, hand-computed split cases
 *
 * Parity unpinned (no IPP here): bit-level rounding of ippsTone_32fc,
 * ippsAbs_32fc_A11, ippsDFT*_32fc and ippsMeanStdDev_32f.  The FP64 variant is
 * the parity oracle; the FP32 variant mirrors the reference's FP32 data path
 * and is the timed CPU baseline.
 */
#ifndef UGSDR_REFCPU_H
#define UGSDR_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes -- */
/* GPS L1 C/A Gold code, sv 0-based (src/prn_codes/GpsL1Ca.hpp:9-21,
 * src/prn_codes/GoldCodes.hpp:24-41).  out: 1023 chips of +-1.  0 = ok. */
int refcpu_gps_l1ca(int sv, int8_t* out);
/* GPS L5 I / Q (src/prn_codes/L5.hpp:25-43, GpsL5I.hpp:9-21, GpsL5Q.hpp:9-21).
 * out: 10230 chips. */
int refcpu_gps_l5i(int sv, int8_t* out);
int refcpu_gps_l5q(int sv, int8_t* out);
/* GLONASS open m-sequence (src/prn_codes/GlonassOf.hpp:24-36). out: 511. */
int refcpu_glonass_of(int8_t* out);
/* BOC(1,1) expansion c -> (c, -c) (src/prn_codes/MemoryCodes.hpp:32-40). */
void refcpu_boc_expand(const int8_t* chips, size_t n, int8_t* out);

/* ------------------------------------------------------------- resample -- */
/* RepeatCodeNTimes (src/common.hpp:78-85) followed by
 * SequentialUpsampler::Process (src/resample/upsampler.hpp:31-37):
 * dst[i] = src[(i*len/samples) % len] with len = n*repeats. */
void refcpu_repeat_upsample(const int8_t* code, size_t n, size_t repeats,
                            float* dst, size_t samples);
/* Real upsampler on floats (src/resample/upsampler.hpp:31-37). */
void refcpu_upsample_f(const float* src, size_t len, float* dst, size_t samples);
/* Sample-and-hold upsampler on complex samples
 * (src/resample/ipp_upsampler.hpp:30-43; ratio = samples/len). */
void refcpu_upsample_hold_c(const float* src, size_t len, float* dst, size_t samples);
/* Accumulator boxcar decimation on complex samples
 * (src/resample/decimator.hpp:53-69, ipp_decimator.hpp:70-88): the tail is
 * zero-padded to a whole block.  Returns the output length. */
size_t refcpu_accumulate_c(const float* src, size_t n, size_t ratio, float* dst);
/* Resampler::Transform (src/resample/resampler.hpp:39-46,
 * ipp_resampler.hpp:131-143): up to lcm(new,old), boxcar down.  dst must hold
 * refcpu_resample_len(); returns the output length. */
size_t refcpu_resample_len(size_t n, size_t new_fs, size_t old_fs);
size_t refcpu_resample_c(const float* src, size_t n, size_t new_fs, size_t old_fs, float* dst);
/* FastSearchEngineBase::AdjustSamplingRate (src/acquisition/fse.hpp:222-240). */
double refcpu_adjust_sampling_rate(double fs, double target);

/* -------------------------------------------------------------- formats -- */
/* Recorded sample formats of SignalParametersBase (src/signal_parameters.hpp:
 * 23-32 FileType).  Pinned by bit-field layouts built independently with
 * ctypes, the reference's Nt1065 first-sample KAT
 * (test/integration_tests.cpp:144-176) and its BbpDdc recordings in data/. */
/* BbpDdc epoch length in bytes at fs (signal_parameters.hpp:74-80). */
int64_t refcpu_bbp_epoch_bytes(double fs);
/* VerifyHeaders (signal_parameters.hpp:36-45): 0 when the 24-bit ms counters
 * of consecutive epochs step by one, -1 otherwise. */
int refcpu_bbp_verify_headers(const uint8_t* data, int64_t n_epochs, int64_t epoch_bytes);
/* GetPartialSignal for BbpDdc from epoch 0 (signal_parameters.hpp:225-236
 * with PackedSpan, src/helpers/BbpPackedSpan.hpp:9-86): n complex samples. */
void refcpu_unpack_bbp(const uint8_t* data, int64_t n, int64_t spm, int64_t epoch_bytes, float* out);
/* NtlabPackedSpan<offset> (src/helpers/NtlabPackedSpan.hpp:9-99): n samples,
 * real value in the real part. */
void refcpu_unpack_nt1065(const uint8_t* data, int64_t n, int offset_bits, float* out);

/* ---------------------------------------------------------------- mixer -- */
/* Mixer::Translate const overload (src/mixer/mixer.hpp:43-51) with the
 * IppMixer tone (src/mixer/ipp_mixer.hpp:83-88): f<0 folds to fs+f, phase is
 * reduced to [0,2pi), x[n] *= exp(j(2 pi f n / fs + phase)).  FP64 tone. */
void refcpu_mix_c(float* x, size_t n, double fs, double f, double phase);
/* In-place overload (src/mixer/mixer.hpp:33-41): returns early for f == 0. */
void refcpu_mix_inplace_c(float* x, size_t n, double fs, double f, double phase);
/* Member Translate (src/mixer/mixer.hpp:53-60): mixes with the stored
 * (fs, f, *phase) and advances *phase by 2 pi frac(f/1000). */
void refcpu_mixer_member_translate_c(float* x, size_t n, double fs, double f, double* phase);
/* TableMixer (src/mixer/table_mixer.hpp:27-101, src/mixer/nco.hpp:8-22), the
 * mixer of the reference's non-IPP configurations: 32-bit phase accumulator
 * with adder = floor(2^32 f/fs) and start floor(2^32 phase/2pi), the top 6
 * bits index an 80-entry float sine table (5/4 period), cos = entry + 16.
 * refcpu_table_mix_c is Mixer::Translate's in-place overload over it
 * (mixer.hpp:33-41), refcpu_table_mixer_member_c the member overload
 * (mixer.hpp:53-60).  Bit-exact with the reference build (oracle/_ref). */
void refcpu_table_mixer_table(float* table80);
void refcpu_table_mix_c(float* x, size_t n, double fs, double f, double phase);
void refcpu_table_mixer_member_c(float* x, size_t n, double fs, double f, double* phase);

/* ------------------------------------------------------------------ fft -- */
/* Arbitrary-N DFT, forward unscaled, inverse scaled by 1/N
 * (src/math/ipp_dft.hpp:39-61, IPP_FFT_DIV_INV_BY_N). data: n complex. */
int refcpu_fft_d(double* data, size_t n, int inverse);
int refcpu_fft_f(float* data, size_t n, int inverse);

/* --------------------------------------------------------- matched filter */
/* IppMatchedFilter::Prepare (src/matched_filter/ipp_matched_filter.hpp:71-78):
 * conj(DFT(real code)). spec: n complex. */
int refcpu_mf_prepare_f(const float* code, size_t n, float* spec);
/* IppMatchedFilter::ProcessOptimized (ipp_matched_filter.hpp:80-90):
 * x <- IDFT(DFT(x) * spec) / n, in place. */
int refcpu_mf_process_f(float* x, const float* spec, size_t n);

/* ----------------------------------------------------------------- math -- */
void refcpu_abs_c(const float* x, size_t n, float* out);                 /* src/math/abs.hpp:20-35 */
int refcpu_reshape_sum_f(const float* x, size_t n, size_t block, float* out); /* src/math/reshape_and_sum.hpp:14-46; -1 on remainder */
void refcpu_max_index_f(const float* x, size_t n, float* value, int64_t* index); /* src/math/max_index.hpp:31-42 (first index) */
void refcpu_mean_std_f(const float* x, size_t n, double* mean, double* sigma); /* src/math/mean_stddev.hpp:29-42 (n-1) */

/* ---------------------------------------------------------- acquisition -- */
typedef struct {
    int64_t n_fft;         /* samples per coherent block at fs_acq              */
    int64_t n_blocks;      /* K non-coherent blocks, block k starts at k*n_fft   */
    int64_t surface_len;   /* fold length: spm (reshape) or n_fft (no reshape)  */
    int32_t reshape;       /* GetOneMsPeak<reshape,...> (src/acquisition/fse.hpp:201-220) */
    int32_t coherent;      /* 1: |.| then fold; 0: fold complex then |.|      */
    double fs_in;          /* signal rate; code_offset = fs_in/fs_acq * index    */
    double fs_acq;         /* acquisition rate                                  */
    double intermediate_frequency; /* added to the reported Doppler          */
    double threshold;      /* 3.3 (src/acquisition/fse.hpp:150)                 */
    int64_t n_doppler;
    const double* doppler; /* Doppler bins in loop order (fse.hpp:250-252)     */
    int64_t n_prn;
    const float* replicas; /* n_prn x n_fft real replica samples               */
} refcpu_acq_params;

typedef struct {
    double level;          /* max of the best Doppler bin's folded surface      */
    double sigma;          /* mean + sample sigma (fse.hpp:261)                 */
    double mean;
    double stddev;
    double doppler;        /* bin frequency + IF (fse.hpp:263)                  */
    double code_offset;    /* ratio * index (fse.hpp:262)                       */
    double snr;            /* level / sigma (acquisition_result.hpp:25-27)      */
    double runner_up_code;    /* second largest surface value in the best row  */
    double runner_up_doppler; /* largest level among the other Doppler bins    */
    int64_t index;
    int32_t doppler_bin;   /* -1 when no bin beat the default result            */
    int32_t detected;      /* snr > threshold (or !reshape, fse.hpp:272)        */
} refcpu_acq_result;

/* FastSearchEngineBase::ProcessBpsk (src/acquisition/fse.hpp:242-276)
 * generalised to K blocks: per PRN, per Doppler bin d (in the given order),
 * per block k: mix the block by -d (tone restarting at phase 0), matched
 * filter against conj(DFT(replica)), |.| (or fold first), fold to
 * surface_len and sum over k; max/first-index, mean and n-1 sigma; keep the
 * bin when strictly greater (fse.hpp:265).  signal: n_blocks*n_fft complex at
 * fs_acq.  precision: 64 = FP64 (parity oracle), 32 = FP32 (reference data
 * path, CPU baseline).  surfaces (optional): n_prn x n_doppler x surface_len.
 * peaks (optional): n_prn x surface_len, the output_peak of the best bin. */
int refcpu_acquire(const refcpu_acq_params* p, const float* signal, int precision,
                   int threads, refcpu_acq_result* out, float* surfaces, float* peaks);

/* ------------------------------------------------------------ correlator -- */
/* SequentialCorrelator::Process (src/correlator/correlator.hpp:28-34):
 * sum signal[i]*code[i], complex x real, FP64 accumulation. out[2]. */
void refcpu_correlate(const float* signal, const float* code, size_t n, double* out);
/* TrackingParameters::CorrelateSplit (src/tracking/tracking_parameters.hpp:471-493)
 * with AddWithPhase (:89-98).  code3: 3-period replica table. */
void refcpu_correlate_split(const float* signal, size_t spm, const float* code3,
                            double code_period_ms, double code_phase, double* out);
/* Tracker::GetEpl (src/tracking/tracker.hpp:119-139) for one block and a tap
 * list: x is mixed in place by (-f_carr, -phase_carr) with the in-place
 * Translate semantics, then CorrelateSplit at code_phase + tap*samples_per_chip.
 * out: n_taps complex (double). */
void refcpu_epl(float* x, size_t spm, double fs, double carrier_frequency, double carrier_phase,
                const float* code3, double code_period_ms, double code_phase,
                double samples_per_chip, const double* taps_chips, int n_taps, double* out);

/* Tracking loop state of one channel (TrackingParameters,
 * src/tracking/tracking_parameters.hpp:100-135); the layout of
 * ugsdr_track_state. */
typedef struct {
    double code_phase, code_frequency, base_code_frequency;
    double carrier_phase, carrier_frequency, carrier_phase_error;
    double code_nco, code_error;
    float previous_prompt[2];
    int64_t epochs;
    int32_t code_period_ms, reserved;
} refcpu_track_state;
/* Tracker::Track for one channel (src/tracking/tracker.hpp:156-177):
 * per epoch b of x (n_blocks x spm complex): the epoch's code phase
 * (code_phase - fs/1e3 * (epochs % code_period)), refcpu_epl at taps
 * {+spacing, 0, -spacing} chips with the carrier phase advanced after the
 * mix (UpdatePhase), the correlations rounded to complex<float>, then
 * TrackingParameters::Pll and ::Dll (tracking_parameters.hpp:495-528) in
 * their float/double mix.  code3: the 3-period replica table.  epl:
 * n_blocks x 3 complex float; trace (optional): n_blocks x 4 doubles
 * (-carrier_phase/2pi, carrier frequency, code phase, code frequency). */
void refcpu_track(const float* x, int64_t spm, int64_t n_blocks, double fs, const float* code3, double spacing,
                  refcpu_track_state* st, float* epl, double* trace);

/* Correlator bank, NCO mode (no reference counterpart; the reference's
 * per-sample replica is an integer shift of the 3-period table).  For each
 * channel c and block b, with block-start carrier phase theta (cycles) and
 * code phase tau (chips), both accumulated in FP64:
 *   s[n]   = x[n] * exp(-j 2 pi (theta + f_c n / fs))
 *   acc[t] = sum_n s[n] * code_c[floor(tau + o_t + n r_c / fs) mod L]
 * then theta += f_c N / fs and tau += r_c N / fs.  x holds n_blocks*spm
 * complex samples; codes n_ch x code_len; out n_blocks x n_ch x n_taps complex. */
typedef struct {
    double carrier_frequency;  /* Hz                         */
    double carrier_phase;      /* cycles at block 0          */
    double code_rate;          /* chips/s                    */
    double code_phase;         /* chips at block 0 sample 0  */
} refcpu_channel;
void refcpu_bank_nco(const float* x, int64_t spm, int64_t n_blocks, double fs,
                     const refcpu_channel* ch, int n_ch, const int8_t* codes, int64_t code_len,
                     const double* taps_chips, int n_taps, int threads, double* out);

#ifdef __cplusplus
}
#endif
#endif
