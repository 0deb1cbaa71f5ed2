/*
 * ugsdr_hip.h — C ABI of the MI355X (gfx950) acquisition / correlator library.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative UGSDR_ERR_* code; ugsdr_hip_last_error() gives the message of the
 * last failure on the calling thread.  No exception crosses this boundary.
 * The caller owns host buffers; the library owns device memory and caches.
 * Calls on one context are serialised internally, so a context may be shared
 * by several host threads (the reference's par_unseq callers).
 *
 * Citations (file:line) name the reference interface each entry replaces,
 * relative to the ugsdr checkout.
 */
#ifndef UGSDR_HIP_H
#define UGSDR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UGSDR_OK 0
#define UGSDR_ERR_ARG (-1)         /* bad argument / null pointer                      */
#define UGSDR_ERR_HIP (-2)         /* HIP runtime failure                              */
#define UGSDR_ERR_SIZE (-3)        /* size contract violated (reference throws)        */
#define UGSDR_ERR_UNSUPPORTED (-4) /* shape not supported by the device path           */
#define UGSDR_ERR_NOMEM (-5)       /* device memory / admission limit                  */
#define UGSDR_ERR_STATE (-6)       /* unknown key, closed server, ...                  */

/* sample formats (src/signal_parameters.hpp:23-32 FileType, IQ subset) */
#define UGSDR_FMT_CI8 1  /* interleaved int8 I/Q   (Iq_8_plus_8)   */
#define UGSDR_FMT_CI16 2 /* interleaved int16 I/Q  (Iq_16_plus_16) */
#define UGSDR_FMT_CF32 3 /* interleaved float I/Q                  */
#define UGSDR_FMT_R8 4   /* real int8 (Real_8), imaginary part 0   */
/* packed recordings (FileType BbpDdc and Nt1065Grabber*; decoders
 * src/helpers/BbpPackedSpan.hpp:9-86 and src/helpers/NtlabPackedSpan.hpp:9-99,
 * used by SignalParametersBase::GetPartialSignal, src/signal_parameters.hpp:140-237) */
#define UGSDR_FMT_BBP_DDC 5  /* epochs: 8-byte header (ms counter) + 2+2-bit complex samples */
#define UGSDR_FMT_NT1065_1 6 /* one byte per sample, channel in bits 6-7 (Nt1065GrabberFirst) */
#define UGSDR_FMT_NT1065_2 7 /* bits 4-5 (Nt1065GrabberSecond) */
#define UGSDR_FMT_NT1065_3 8 /* bits 2-3 (Nt1065GrabberThird)  */
#define UGSDR_FMT_NT1065_4 9 /* bits 0-1 (Nt1065GrabberFourth) */

/* arithmetic of the correlation surfaces */
#define UGSDR_PREC_FP32 0 /* plain: FP32 complex like the reference            */
#define UGSDR_PREC_FP16 1 /* reduced: FP16 storage of samples and spectra      */
#define UGSDR_PREC_BF16 2 /* reduced: BF16 storage of samples and spectra      */

typedef struct ugsdr_hip_ctx ugsdr_hip_ctx;

const char* ugsdr_hip_last_error(void);
/* library version string */
const char* ugsdr_hip_version(void);

/* ------------------------------------------------------------------ context */
/* Creates a context on `device`.  Replaces the implicit per-thread state of
 * the reference's operators (thread_local scratch in src/mixer/ipp_mixer.hpp:15-19,
 * src/math/ipp_dft.hpp:24-41) with one device-resident owner. */
int ugsdr_hip_ctx_create(int device, ugsdr_hip_ctx** out);
int ugsdr_hip_ctx_destroy(ugsdr_hip_ctx* ctx);
/* the HIP stream (hipStream_t) the context launches on, for event timing */
int ugsdr_hip_ctx_stream(ugsdr_hip_ctx* ctx, void** stream);
/* bytes of device memory held by the context (caches + scratch) */
int64_t ugsdr_hip_ctx_device_bytes(ugsdr_hip_ctx* ctx);
/* Stage timing with HIP events on the context's stream (off by default).  No
 * reference counterpart: the reference times nothing at run time; this is the
 * harness hook that replaces its disabled google-benchmark section
 * (benchmark/benchmark.cpp:482-539).  ms_out[4] = front end (unpack, mix,
 * resample), forward class FFTs, correlation, finalize of the last acquire;
 * after ugsdr_hip_track or ugsdr_hip_correlate_bank, ms_out[2] is that call's
 * kernel and the others are 0. */
int ugsdr_hip_ctx_set_profiling(ugsdr_hip_ctx* ctx, int enable);
int ugsdr_hip_ctx_last_times(ugsdr_hip_ctx* ctx, float* ms_out4);

/* Autotuner (no reference counterpart): times the correlation kernel's
 * workgroup-size variants for a transform length on a synthetic 32 PRN x 32
 * bin batch and keeps the fastest for this context.  Runs by itself on the
 * first acquisition of each length unless disabled.  threads: chosen size;
 * ms_out[n_ms]: each candidate's time per launch. */
int ugsdr_hip_autotune(ugsdr_hip_ctx* ctx, int64_t n_fft, int32_t* threads, float* ms_out, int32_t n_ms);
int ugsdr_hip_set_autotune(ugsdr_hip_ctx* ctx, int enable);
/* hipGraph replay of repeated searches (no reference counterpart; on by
 * default): when ugsdr_hip_acquire sees the same request shape (everything
 * but the host IQ pointer and its contents) twice in a row, it captures the
 * search's kernels into a graph and replays it on later calls; the host IQ
 * upload and the result copies stay outside the graph.  Results are those
 * of the eager path.  graph_replays counts the replays of this context. */
int ugsdr_hip_set_graphs(ugsdr_hip_ctx* ctx, int enable);
int64_t ugsdr_hip_graph_replays(ugsdr_hip_ctx* ctx);
/* The form the last ugsdr_hip_acquire on this context took (no reference
 * counterpart; the results do not depend on it): n_transform, the points of
 * each inverse transform's row (n_fft, or the surface length of a
 * periodic-fold search); periods, the code periods combined per block by the
 * periodic fold (1: none); samples_per_chip, the upsampled-code search's
 * factor (0: not taken; the row then runs as samples_per_chip inverse
 * transforms of n_transform / samples_per_chip points on spectra shared by
 * the PRNs of a Doppler bin).  Any pointer may be null. */
int ugsdr_hip_last_search_form(ugsdr_hip_ctx* ctx, int64_t* n_transform, int32_t* periods, int32_t* samples_per_chip);
/* The correlation variants built for a transform length, each named by its
 * workgroup size (a 1023*M length also has the three-stage kernel, named by
 * its own size: 192 at N = 4092, 512 at N = 8184, 12276 and 16368, where it
 * replaces the generic kernel of that size).  Writes up to n of them to
 * threads_out and returns how many exist (no reference counterpart). */
int ugsdr_hip_corr_variants(int64_t n_fft, int32_t* threads_out, int32_t n);
/* Pins the correlation variant for a length on this context (threads from
 * ugsdr_hip_corr_variants; 0 unpins, returning to the autotuner/default). */
int ugsdr_hip_set_threads(ugsdr_hip_ctx* ctx, int64_t n_fft, int32_t threads);

/* ------------------------------------------------------------- PRN codes (host)
 * Codegen<Impl>::Get (src/prn_codes/codegen.hpp:23-36).  sv is 0-based. */
int ugsdr_code_gps_l1ca(int sv, int8_t* out1023);  /* GpsL1Ca.hpp:9-21, GoldCodes.hpp:24-41 */
int ugsdr_code_gps_l5i(int sv, int8_t* out10230);  /* GpsL5I.hpp:9-21, L5.hpp:25-43 */
int ugsdr_code_gps_l5q(int sv, int8_t* out10230);  /* GpsL5Q.hpp:9-21, L5.hpp:25-43 */
int ugsdr_code_glonass_of(int8_t* out511);         /* GlonassOf.hpp:24-36 */
/* BOC(1,1) chip expansion c -> (c,-c): MemoryCodes::DecodeBoc (MemoryCodes.hpp:32-40) */
int ugsdr_code_boc(const int8_t* chips, int64_t n, int8_t* out2n);
/* FastSearchEngineBase::AdjustSamplingRate (src/acquisition/fse.hpp:222-240) */
double ugsdr_adjust_sampling_rate(double fs, double target);

/* ------------------------------------------------------------ replica cache
 * Device-resident conj(DFT(replica)) of length n_fft.  Replaces
 * MatchedFilter::PrepareCodeSpectrum (src/matched_filter/matched_filter.hpp:27-30,
 * ipp_matched_filter.hpp:71-78) applied to
 * Upsampler::Transform(RepeatCodeNTimes(code, repeats), n_fft)
 * (src/acquisition/fse.hpp:295-296, src/common.hpp:78-85,
 * src/resample/upsampler.hpp:31-37): replica[i] = chips[(i*len/n_fft) % len],
 * len = n_chips*repeats.  `precision` is UGSDR_PREC_*. */
int ugsdr_hip_replica_put(ugsdr_hip_ctx* ctx, const int8_t* chips, int64_t n_chips, int64_t repeats,
                          int64_t n_fft, int precision, uint64_t* key);
/* same from an arbitrary real replica of n_fft samples (e.g. the per-ms sign
 * permutations of src/acquisition/fse.hpp:374-385) */
int ugsdr_hip_replica_put_samples(ugsdr_hip_ctx* ctx, const float* samples, int64_t n_fft, int precision,
                                  uint64_t* key);
/* frees a cached spectrum; UGSDR_ERR_STATE while queued or running server
 * requests read it (ugsdr_hip_server_submit pins the request's keys until its
 * result is ready) */
int ugsdr_hip_replica_drop(ugsdr_hip_ctx* ctx, uint64_t key);
/* copies the stored spectrum (n_fft complex float) back, for inspection */
int ugsdr_hip_replica_get(ugsdr_hip_ctx* ctx, uint64_t key, float* spectrum_c32, int64_t n_fft);

/* ------------------------------------------------------- fused acquisition */
typedef struct {
    const void* iq;        /* samples at fs_in in `fmt` (BBP_DDC: from an epoch header) */
    int32_t fmt;           /* UGSDR_FMT_*                                            */
    int32_t iq_on_device;  /* 1: iq is a device pointer, 0: host pointer             */
    int64_t n_samples;     /* complex samples available at iq                        */
    double fs_in;          /* input rate                                             */
    double fs_acq;         /* acquisition rate (Resampler::Transform target)         */
    double mix_frequency;  /* applied first as Mixer::Translate(x, fs_in, f, phase)  */
    double mix_phase;      /*   (the -IF wipe-off of fse.hpp:289, and/or a DFE tone)  */
    double intermediate_frequency; /* added to reported Doppler (fse.hpp:263)       */
    int64_t n_fft;         /* samples per coherent block at fs_acq                   */
    int64_t n_blocks;      /* K non-coherent blocks                                  */
    int64_t surface_len;   /* fold length: spm (reshape) or n_fft (no reshape)       */
    int32_t reshape;       /* GetOneMsPeak<reshape, coherent> (fse.hpp:201-220)      */
    int32_t coherent;
    const double* doppler; /* host array, loop order of fse.hpp:250-252              */
    int64_t n_doppler;
    const uint64_t* replica_keys; /* host array, one per PRN, each of length n_fft */
    int64_t n_prn;
    double threshold;      /* 3.3 in the reference (fse.hpp:150)                     */
    int32_t precision;     /* UGSDR_PREC_*                                           */
    int32_t hypotheses;    /* 0 or 1: one result per replica.  H > 1: each run of H
                            * consecutive replicas is one satellite's code
                            * hypotheses (the Galileo E1 sign permutations of
                            * fse.hpp:374-389): one result per run, the hypothesis
                            * of largest level (the first on ties, max_element),
                            * detected when its SNR clears the threshold.  n_prn
                            * must be a multiple of H; surfaces and peaks stay per
                            * replica.                                              */
} ugsdr_acq_request;

typedef struct {
    double level;        /* max of the best bin's surface (reference scale)  */
    double sigma;        /* mean + sample sigma (fse.hpp:261)                */
    double mean;
    double stddev;
    double doppler;      /* bin frequency + IF                               */
    double code_offset;  /* fs_in/fs_acq * index (fse.hpp:262)               */
    double snr;          /* level / sigma                                    */
    double runner_up_code;    /* second value of the best row              */
    double runner_up_doppler; /* best level of the other Doppler bins      */
    int64_t index;
    int32_t doppler_bin; /* -1 if no bin beat the default (level 0)          */
    int32_t detected;    /* snr > threshold, or !reshape (fse.hpp:272)       */
} ugsdr_acq_result;

/* FastSearchEngineBase::ProcessBpsk for n_prn PRNs at once
 * (src/acquisition/fse.hpp:242-276, generalised to K blocks): mix, resample,
 * per (PRN, Doppler, block) circular correlation against the cached replica
 * spectrum, |.| and fold, sum over blocks, first-index max, mean, n-1 sigma,
 * strict-< Doppler selection, threshold.  out: n_prn results (n_prn / H with
 * hypotheses H > 1: FastSearchEngineBase::ProcessGalileo's selection,
 * fse.hpp:356-396).  surfaces
 * (optional host buffer): n_prn x n_doppler x surface_len.  peaks (optional):
 * n_prn x surface_len, the output_peak of the selected bin. */
int ugsdr_hip_acquire(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* req, ugsdr_acq_result* out, float* surfaces,
                      float* peaks);

/* The same search for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Arguments and results are
 * those of ugsdr_hip_acquire for the FP32 plain form (UGSDR_ERR_UNSUPPORTED
 * for reduced precision).  Every device buffer of the search is carved out of
 * one fixed 64 MiB device array per GPU, held until the process ends; the
 * call runs no hipMalloc / hipFree, graph instantiation or autotuning (it uses
 * the context's tuned workgroup size if one exists, else the default).  A
 * search whose buffers exceed the array (e.g. with full surfaces of a large
 * search) returns UGSDR_ERR_NOMEM.  Calls on one GPU serialize on a lock
 * inside the library because they share its array, and the array keeps the
 * last call's contents (no kernel reads what it has not written in the same
 * call).  Host-side bookkeeping of the request (its plan, a few KB) uses the
 * ordinary host heap. */
int ugsdr_hip_acquire_noalloc(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* req, ugsdr_acq_result* out,
                              float* surfaces, float* peaks);

/* ------------------------------------------------------ anti-jamming (STFT)
 * The interference suppression of the digital front end
 * (src/antijamming/jse.hpp, jamming_detection.hpp, src/math/stft.hpp). */
#define UGSDR_STFT_BLACKMAN_HARRIS 0 /* IppStft: 4-term Blackman-Harris over N-1 (ipp_stft.hpp:26-45) */
#define UGSDR_STFT_HANN 1            /* SequentialStft: periodic Hann (stft.hpp:31-37)                 */
#define UGSDR_JAMMER_NARROWBAND 1
#define UGSDR_JAMMER_WIDEBAND 2
typedef struct {
    int32_t type;           /* UGSDR_JAMMER_* (DetectionResult::InterferenceType) */
    int32_t frequency_bin;  /* STFT bin, 0..127                                    */
    double frequency;       /* bin * fs / 128                                      */
    double jammer_to_noise; /* mean J/N over the columns that voted for it         */
} ugsdr_jammer;
/* ShortTimeFourierTransform::Transform (stft.hpp:16-20,40-58): (n - 96) / 32
 * columns of 128 windowed samples at stride 32, forward DFT without scaling.
 * out: n_cols x 128 complex float (host buffers). */
int ugsdr_hip_stft(ugsdr_hip_ctx* ctx, const float* in_c32, int64_t n, int window, float* out_c32);
/* the container-of-columns overload, OverlapAndAdd (stft.hpp:60-72): inverse
 * DFT (1/128) of each column added at stride 32; out: 128 + 32 (n_cols - 1)
 * complex float */
int ugsdr_hip_istft(ugsdr_hip_ctx* ctx, const float* cols_c32, int64_t n_cols, float* out_c32);
/* JammingSuppressionEngine::Process (jse.hpp:93-103) in place on n complex
 * float samples at fs (host, or device memory with on_device): STFT
 * detection (JammingDetector::Process), then nothing, the whole-signal
 * notch of +-128 DFT bins per narrowband detection, or the STFT excision of
 * cells with |X| >= 20 when any detection is wideband.  Up to max_out
 * detections go to out; *n_out gets their count.  The notch's whole-signal
 * DFT takes lengths whose prime factors are at most 4096 (UGSDR_ERR_UNSUPPORTED
 * otherwise, samples untouched). */
int ugsdr_hip_jse_process(ugsdr_hip_ctx* ctx, float* io_c32, int64_t n, double fs, int window, int on_device,
                          ugsdr_jammer* out, int32_t max_out, int32_t* n_out);

/* ------------------------------------------------------ recorded formats
 * Bytes that hold n samples of fmt at fs (BBP_DDC: whole epochs, headers
 * included, epoch length as in the file geometry of
 * src/signal_parameters.hpp:74-80; fs must be a whole number of kHz), or a
 * negative UGSDR_ERR_* code. */
int64_t ugsdr_format_bytes(int fmt, int64_t n_samples, double fs);
/* SignalParametersBase::GetPartialSignal from the start of a recording
 * (src/signal_parameters.hpp:140-237): decodes n samples of fmt into complex
 * float (host buffers).  BBP_DDC first checks the epoch ms counters like
 * VerifyHeaders (signal_parameters.hpp:36-45; UGSDR_ERR_ARG on a gap). */
int ugsdr_hip_unpack(ugsdr_hip_ctx* ctx, const void* in, int fmt, int64_t n, double fs, float* out_c32);

/* ------------------------------------------------------ DFE channel step
 * One digital-front-end channel (Channel, src/dfe/dfe.hpp:110-203): recorded
 * samples in any UGSDR_FMT_* at fs_in are decoded, mixed by mixer_frequency
 * (the file's centre frequency minus the signal's, dfe.hpp:177-180) with
 * the channel's running phase, and resampled to fs_out.  The interference
 * suppression option (dfe.hpp:200-201) is not part of this step. */
typedef struct ugsdr_hip_dfe ugsdr_hip_dfe;
/* fs_in and fs_out must be whole numbers of kHz (epochs of 1 ms). */
int ugsdr_hip_dfe_create(ugsdr_hip_ctx* ctx, int fmt, double fs_in, double fs_out, double mixer_frequency,
                         double phase, ugsdr_hip_dfe** out);
/* Channel::GetSeveralEpochs (dfe.hpp:192-203) without suppression: n_epochs
 * ms of the recording at `in` (from an epoch boundary; host or device
 * memory) through Mixer::Translate's member overload (mixer.hpp:53-60: the
 * stored phase, then phase += 2 pi frac(f / 1000) per call) and
 * Resampler::Transform (resampler.hpp:39-46).  out: n_epochs * fs_out/1000
 * complex samples, host memory or, with out_on_device = 1, device memory
 * (e.g. the iq of a later ugsdr_hip_acquire with iq_on_device = 1; it is
 * complete when the call returns). */
int ugsdr_hip_dfe_process(ugsdr_hip_dfe* ch, const void* in, int in_on_device, int64_t n_epochs, float* out_c32,
                          int out_on_device);
/* the channel's running mixer phase (radians, not reduced) */
int ugsdr_hip_dfe_phase(ugsdr_hip_dfe* ch, double* phase);
int ugsdr_hip_dfe_destroy(ugsdr_hip_dfe* ch);

/* ---------------------------------------- per-operator drop-ins (host data)
 * Slow path behind the Hip* CRTP classes: one call = one reference operator
 * call, data copied in and out like the ArrayFire proxy backends. */
/* Mixer::Translate const overload + IppMixer::Process (src/mixer/mixer.hpp:43-51,
 * ipp_mixer.hpp:83-88): f<0 folds to fs+f, phase reduced to [0,2pi). */
int ugsdr_hip_mix(ugsdr_hip_ctx* ctx, const float* in_c32, float* out_c32, int64_t n, double fs, double f,
                  double phase);
/* Mixer::Translate in-place overload (mixer.hpp:33-41): no-op when f == 0 */
int ugsdr_hip_mix_inplace(ugsdr_hip_ctx* ctx, float* io_c32, int64_t n, double fs, double f, double phase);
/* Mixer::Translate in-place overload (mixer.hpp:33-41) over TableMixer, the
 * mixer of the reference's non-IPP configurations (src/mixer/table_mixer.hpp:
 * 27-105, src/mixer/nco.hpp:8-22): a 32-bit phase accumulator started at
 * floor(2^32 phase / 2 pi) and advanced by floor(2^32 f / fs) per sample, whose
 * top 6 bits index an 80-entry float sine table (cos 16 entries on).  No-op
 * when f == 0; f < 0 folds to fs + f (UGSDR_ERR_ARG below -fs).  io holds n
 * complex float samples in host memory, or device memory with on_device = 1
 * (complete when the call returns).  Bit-exact with the reference build. */
int ugsdr_hip_table_mix(ugsdr_hip_ctx* ctx, float* io_c32, int64_t n, double fs, double f, double phase,
                        int on_device);
/* DiscreteFourierTransform::Transform (src/math/ipp_dft.hpp:106-122): any n,
 * forward unscaled, inverse scaled 1/n */
int ugsdr_hip_dft(ugsdr_hip_ctx* ctx, const float* in_c32, float* out_c32, int64_t n, int inverse);
/* The same transform in reduced precision over a batch of sequences: 16-bit
 * (FP16 or BF16) operands on the matrix cores with FP32 accumulation, for the
 * two-stage lengths of the reduced acquisition path (60, 112, 120, 180, 240,
 * 341).  Forward unscaled, inverse 1/n, as IppDft (ipp_dft.hpp:106-122); the
 * 16-bit form itself has no reference counterpart. */
int ugsdr_hip_dft_reduced(ugsdr_hip_ctx* ctx, int precision, int inverse, const float* in_c32, float* out_c32,
                          int64_t n, int64_t batch);
/* MatchedFilter::PrepareCodeSpectrum (matched_filter.hpp:27-30): conj(DFT(real code)) */
int ugsdr_hip_mf_prepare(ugsdr_hip_ctx* ctx, const float* code, float* spec_c32, int64_t n);
/* MatchedFilter::FilterOptimized (matched_filter.hpp:32-42, ipp_matched_filter.hpp:80-90) */
int ugsdr_hip_mf_filter(ugsdr_hip_ctx* ctx, const float* in_c32, const float* spec_c32, float* out_c32,
                        int64_t n);
/* MatchedFilter::Filter (matched_filter.hpp:15-25): prepare + filter */
int ugsdr_hip_mf_filter_code(ugsdr_hip_ctx* ctx, const float* in_c32, const float* code, float* out_c32,
                             int64_t n);
/* Abs::Transform (src/math/abs.hpp:14-19) */
int ugsdr_hip_abs(ugsdr_hip_ctx* ctx, const float* in_c32, float* out, int64_t n);
/* ReshapeAndSum::Transform (src/math/reshape_and_sum.hpp:27-31); real (complex=0)
 * or complex (complex=1); UGSDR_ERR_SIZE when n % block != 0 */
int ugsdr_hip_reshape_sum(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n, int64_t block, int complex_);
/* MaxIndex::Transform (src/math/max_index.hpp:21-24): first index of the max */
int ugsdr_hip_max_index(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* value, int64_t* index);
/* MeanStdDev::Calculate (src/math/mean_stddev.hpp:21-24): mean, sample sigma */
int ugsdr_hip_mean_std(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* mean, float* sigma);
/* Resampler::Transform (src/resample/resampler.hpp:15-25, ipp_resampler.hpp:131-143):
 * sample-and-hold to lcm, boxcar down.  out must hold ugsdr_resample_len(). */
int64_t ugsdr_resample_len(int64_t n, int64_t new_fs, int64_t old_fs);
int ugsdr_hip_resample(ugsdr_hip_ctx* ctx, const float* in_c32, int64_t n, int64_t new_fs, int64_t old_fs,
                       float* out_c32);
/* Upsampler::Transform (src/resample/upsampler.hpp:13-23, 31-37): real */
int ugsdr_hip_upsample(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* out, int64_t samples);
/* Correlator::Correlate (src/correlator/correlator.hpp:16-20, ipp_correlator.hpp:25-38):
 * sum signal[i]*code[i]; UGSDR_ERR_SIZE on n_signal != n_code */
int ugsdr_hip_correlate(ugsdr_hip_ctx* ctx, const float* signal_c32, int64_t n_signal, const float* code,
                        int64_t n_code, float* out_c32);

/* ------------------------------------------------------- correlator bank */
typedef struct {
    double carrier_frequency;  /* Hz                                             */
    double carrier_phase;      /* ref_split: radians; nco: cycles                */
    double code_rate;          /* nco: chips/s                                   */
    double code_phase;         /* ref_split: samples; nco: chips at block start  */
} ugsdr_channel_state;

#define UGSDR_BANK_REF_SPLIT 0 /* Tracker::GetEpl + CorrelateSplit semantics */
#define UGSDR_BANK_NCO 1       /* FP64 carrier and code NCOs per sample      */

typedef struct {
    const void* iq;            /* n_blocks * spm samples in fmt               */
    int32_t fmt;
    int32_t iq_on_device;
    int64_t spm;               /* samples per block (1 ms)                    */
    int64_t n_blocks;
    double fs;
    int32_t mode;              /* UGSDR_BANK_*                                 */
    int32_t n_channels;
    int32_t n_taps;
    int32_t reserved;
    const double* taps_chips;  /* n_taps tap offsets in chips                  */
    const int8_t* codes;       /* n_channels x code_len chips (+-1), host      */
    int64_t code_len;          /* chips per code period                        */
    double samples_per_chip;   /* ref_split: replica table rate (fs/chip rate) */
    double code_period_ms;     /* ref_split: code period                       */
} ugsdr_bank_request;

/* Correlator bank over n_blocks consecutive blocks.  ch holds each channel's
 * state at block 0 and is advanced in place (carrier phase += 2 pi f/1000 per
 * block in ref_split mode, TrackingParameters::UpdatePhase,
 * src/tracking/tracking_parameters.hpp:163-165; FP64 NCO advance in nco mode).
 * ref_split replaces Tracker::GetEpl (src/tracking/tracker.hpp:119-139) with
 * TrackingParameters::CorrelateSplit (tracking_parameters.hpp:471-493).
 * out: n_blocks x n_channels x n_taps complex float (host buffer).  At most
 * 65535 blocks per call (UGSDR_ERR_SIZE beyond). */
int ugsdr_hip_correlate_bank(ugsdr_hip_ctx* ctx, const ugsdr_bank_request* req, ugsdr_channel_state* ch,
                             float* out_c32);

/* ------------------------------------------------------- tracking loops
 * GPU-resident Tracker::Track (src/tracking/tracker.hpp:156-197): per channel
 * and 1 ms epoch, the epoch is mixed by the carrier NCO, early / prompt /
 * late are correlated with TrackingParameters::CorrelateSplit against the
 * 3-period replica table (tracking_parameters.hpp:471-493, tracker.hpp:
 * 119-139), and the PLL/FLL and DLL updates (TrackingParameters::Pll / Dll,
 * tracking_parameters.hpp:495-528) run on the device before the next epoch.
 * One call runs every channel over n_blocks consecutive epochs. */
typedef struct {
    double code_phase;           /* samples (TrackingParameters::code_phase)     */
    double code_frequency;       /* chips/s                                      */
    double base_code_frequency;  /* chips/s (1.023e6 for GPS L1 C/A)            */
    double carrier_phase;        /* radians                                      */
    double carrier_frequency;    /* Hz                                           */
    double carrier_phase_error;  /* last PLL discriminator, cycles               */
    double code_nco;
    double code_error;           /* last DLL discriminator                       */
    float previous_prompt[2];    /* I, Q                                         */
    int64_t epochs;              /* prompts produced so far (prompt.size())      */
    int32_t code_period_ms;      /* GetCodePeriod (tracking_parameters.hpp:166-199) */
    int32_t reserved;
} ugsdr_track_state;

typedef struct {
    const void* iq;            /* n_blocks * spm samples in fmt (the channel's subband) */
    int32_t fmt;
    int32_t iq_on_device;
    int64_t spm;               /* samples per 1 ms epoch                                */
    int64_t n_blocks;
    double fs;
    int32_t n_channels;
    int32_t reserved;
    const int8_t* codes;       /* n_channels x code_len chips (+-1), host               */
    int64_t code_len;          /* chips per code period                                 */
    double spacing_chips;      /* early/late offset, 0.25 in Tracker (tracker.hpp:170)  */
} ugsdr_track_request;

/* Runs the loops and advances ch (n_channels states) in place.  epl:
 * n_blocks x n_channels x 3 complex (early, prompt, late), host buffer.
 * trace (optional, host): n_blocks x n_channels x 4 doubles, the values
 * TrackingParameters records per epoch: -carrier_phase / 2 pi, carrier
 * frequency, code phase, code frequency (phases, frequencies, code_phases,
 * code_frequencies). */
int ugsdr_hip_track(ugsdr_hip_ctx* ctx, const ugsdr_track_request* req, ugsdr_track_state* ch, float* epl_c32,
                    double* trace);

/* ---------------------------------------------------------- request server
 * Continuous-batching acquisition server (no reference counterpart: the
 * reference runs one FastSearchEngineBase::Process per epoch,
 * src/acquisition/fse.hpp:457-507).  Requests of uneven shape are submitted as
 * they arrive; a scheduler thread admits them while request slots and the
 * device-memory budget allow, runs every admitted request whose transform
 * length matches in one fused correlation launch, and rotates successive
 * rounds over several HIP streams so their work overlaps on the device.
 * Results are per request identical to ugsdr_hip_acquire on the same request. */
#define UGSDR_PENDING 1  /* ugsdr_hip_server_poll: not finished yet */

typedef struct ugsdr_hip_server ugsdr_hip_server;

typedef struct {
    int64_t max_device_bytes;  /* admission budget of in-flight requests (0: 8 GiB)   */
    int32_t max_inflight;      /* request slots (0: 64)                               */
    int32_t n_streams;         /* streams the rounds rotate over (0: 4)               */
    int32_t max_batch;         /* requests per correlation launch (0: 16)             */
    int32_t round_budget_us;   /* batch-shape autotuner: a round of one bucket is cut
                                * when its estimated time (its row-blocks P x D x K
                                * times the bucket's measured time per row-block,
                                * averaged over the bucket's completed rounds) would
                                * exceed this (0: 2000 us); before a bucket's first
                                * round completes, its rounds hold at most 4        */
    int32_t round_graphs;      /* nonzero: each round's device steps are captured
                                * into a hipGraph and run as one launch, replaying
                                * (with updated parameters) an executable of the
                                * same kernel sequence; 0: the steps are launched
                                * one by one, which costs the scheduler about half
                                * the host time per round on ROCm 7.2             */
} ugsdr_server_config;

typedef struct {
    int64_t submitted, completed, failed, inflight, inflight_bytes, rounds, batches, max_batch_seen;
    /* round graphs: launches, executables instantiated, executables updated and
     * replayed, rounds that ran without a graph (no idle executable of the
     * round's kernel sequence yet, or the capture failed) */
    int64_t graph_launches, graph_instantiations, graph_updates, graph_fallbacks;
    int64_t launch_host_us;    /* scheduler time spent launching rounds        */
} ugsdr_server_stats;

int ugsdr_hip_server_create(ugsdr_hip_ctx* ctx, const ugsdr_server_config* cfg, ugsdr_hip_server** out);
/* Validates and copies the request (Doppler list, keys and host IQ), so the
 * caller may reuse its buffers at once; a device IQ pointer must stay valid
 * until the ticket completes.  The request's replica keys stay pinned
 * (ugsdr_hip_replica_drop refuses them) until its result is ready. */
int ugsdr_hip_server_submit(ugsdr_hip_server* srv, const ugsdr_acq_request* req, uint64_t* ticket);
/* UGSDR_OK: finished, out[n_prn] filled, latency (submit to completion) in ms,
 * ticket released; UGSDR_PENDING: still running after wait_ms; < 0: the
 * request failed (message in ugsdr_hip_last_error), ticket released. */
int ugsdr_hip_server_poll(ugsdr_hip_server* srv, uint64_t ticket, int wait_ms, ugsdr_acq_result* out,
                          double* latency_ms);
int ugsdr_hip_server_stats(ugsdr_hip_server* srv, ugsdr_server_stats* stats);
/* drains in-flight work, then frees the server */
int ugsdr_hip_server_destroy(ugsdr_hip_server* srv);

#ifdef __cplusplus
}
#endif
#endif
