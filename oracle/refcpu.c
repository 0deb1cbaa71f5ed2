/*
 * refcpu — CPU restatement of ugsdr's acquisition hot path (oracle).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg; never by the product library.  Declarations and the list
 * of pins are in refcpu.h.
 */
#define _GNU_SOURCE
#include "refcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define REAL double
#define SUFFIX _d
#include "refcpu_fft.inc"
#undef REAL
#undef SUFFIX
#define REAL float
#define SUFFIX _f
#include "refcpu_fft.inc"
#undef REAL
#undef SUFFIX

/* ================================================================ codes == */
/* LfsrBase<27> (src/prn_codes/lfsr.hpp:7-161) */
typedef struct {
    uint32_t value, poly, inv, outpin, reset_period, reset_value, shift_counter, seq_len;
} rc_lfsr;

static const uint32_t RC_LFSR_MASK = 0xFFFFFFFFu >> (32 - 27);

static uint32_t rc_count_triggers(uint32_t poly) {
    for (uint32_t i = 0; i < 27; ++i) {
        if (!poly) return i;
        poly >>= 1;
    }
    return 27;
}

static void rc_lfsr_init(rc_lfsr* l, uint32_t val, uint32_t poly, uint32_t outpin) {
    memset(l, 0, sizeof(*l));
    l->value = val & RC_LFSR_MASK;                       /* Val()  lfsr.hpp:32-34 */
    l->poly = poly & RC_LFSR_MASK;                       /* Poly() lfsr.hpp:41-46 */
    l->seq_len = (1u << rc_count_triggers(l->poly)) - 1;
    l->outpin = outpin;
}

static void rc_lfsr_reset_period(rc_lfsr* l, uint32_t period, uint32_t val) {
    uint32_t trig = rc_count_triggers(l->poly);
    l->reset_period = period;                            /* lfsr.hpp:67-70 */
    l->seq_len = period ? period : ((1u << trig) - 1);
    l->reset_value = val & RC_LFSR_MASK;                 /* lfsr.hpp:77-79 */
}

static int rc_lfsr_get(const rc_lfsr* l) { return (int)((l->value >> l->outpin) & 1u); }

static void rc_lfsr_shift(rc_lfsr* l) {                  /* lfsr.hpp:94-110 */
    ++l->shift_counter;
    if (l->reset_period && l->shift_counter == l->reset_period) {
        l->shift_counter = 0;
        l->value = l->reset_value;
        return;
    }
    uint32_t v = l->value & l->poly;
    uint32_t s = (uint32_t)__builtin_popcount(v);
    l->value = ((l->value << 1) | ((s & 1u) ^ l->inv)) & RC_LFSR_MASK;
}

/* ICD-GPS-200 G2 delay table as listed in src/prn_codes/GpsL1Ca.hpp:10-15 */
static const int rc_l1ca_delay[37] = {
    5, 6, 7, 8, 17, 18, 139, 140, 141, 251, 252, 254, 255, 256, 257, 258, 469, 470, 471,
    472, 473, 474, 509, 512, 513, 514, 515, 516, 859, 860, 861, 862, 863, 950, 947, 948, 950};

int refcpu_gps_l1ca(int sv, int8_t* out) {
    if (sv < 0 || sv >= 37) return -1;                   /* GpsL1Ca.hpp:17-18 */
    rc_lfsr g[2];
    rc_lfsr_init(&g[0], 0x3FF, 0x204, 9);                /* GoldCodes.hpp:10-13 */
    rc_lfsr_init(&g[1], 0x3FF, 0x3A6, 9);
    for (int i = 0; i < 1023 - rc_l1ca_delay[sv]; ++i) rc_lfsr_shift(&g[1]);
    for (uint32_t i = 0; i < g[0].seq_len; ++i) {
        out[i] = (int8_t)(1 - 2 * (rc_lfsr_get(&g[0]) ^ rc_lfsr_get(&g[1])));
        rc_lfsr_shift(&g[0]);
        rc_lfsr_shift(&g[1]);
    }
    return 0;
}

/* XB advances, src/prn_codes/GpsL5I.hpp:10-15 and GpsL5Q.hpp:10-15 */
static const int rc_l5i_adv[32] = {266, 365, 804, 1138, 1509, 1559, 1756, 2084, 2170, 2303, 2527,
                                   2687, 2930, 3471, 3940, 4132, 4332, 4924, 5343, 5443, 5641, 5816,
                                   5898, 5918, 5955, 6243, 6345, 6477, 6518, 6875, 7168, 7187};
static const int rc_l5q_adv[32] = {1701, 323, 5292, 2020, 5429, 7136, 1041, 5947, 4315, 148, 535,
                                   1939, 5206, 5910, 3595, 5135, 6082, 6990, 3546, 1523, 4548, 4484,
                                   1893, 3961, 7106, 5299, 4660, 276, 4389, 3783, 1591, 1601};

static int rc_l5(int adv, int8_t* out) {                 /* src/prn_codes/L5.hpp:25-43 */
    rc_lfsr g[2];
    rc_lfsr_init(&g[0], 0xFFFF, 0x1B00, 12);
    rc_lfsr_init(&g[1], 0xFFFF, 0x18ED, 12);
    rc_lfsr_reset_period(&g[0], 8190, 0xFFFF);
    for (int i = 0; i < adv; ++i) rc_lfsr_shift(&g[1]);
    for (int i = 0; i < 10230; ++i) {
        out[i] = (int8_t)(1 - 2 * (rc_lfsr_get(&g[0]) ^ rc_lfsr_get(&g[1])));
        rc_lfsr_shift(&g[0]);
        rc_lfsr_shift(&g[1]);
    }
    return 0;
}

int refcpu_gps_l5i(int sv, int8_t* out) { return (sv < 0 || sv >= 32) ? -1 : rc_l5(rc_l5i_adv[sv], out); }
int refcpu_gps_l5q(int sv, int8_t* out) { return (sv < 0 || sv >= 32) ? -1 : rc_l5(rc_l5q_adv[sv], out); }

int refcpu_glonass_of(int8_t* out) {                     /* GlonassOf.hpp:24-36 */
    rc_lfsr l;
    rc_lfsr_init(&l, 0x1FF, 0x110, 6);
    for (uint32_t i = 0; i < l.seq_len; ++i) {
        out[i] = (int8_t)(1 - 2 * rc_lfsr_get(&l));
        rc_lfsr_shift(&l);
    }
    return 0;
}

void refcpu_boc_expand(const int8_t* chips, size_t n, int8_t* out) {
    for (size_t i = 0; i < n; ++i) { out[2 * i] = chips[i]; out[2 * i + 1] = (int8_t)(-chips[i]); }
}

/* ============================================================= resample == */
void refcpu_repeat_upsample(const int8_t* code, size_t n, size_t repeats, float* dst, size_t samples) {
    const size_t len = n * repeats;                      /* common.hpp:78-85 */
    for (size_t i = 0; i < samples; ++i)                 /* upsampler.hpp:31-37 */
        dst[i] = (float)code[((i * len / samples) % len) % n];
}

void refcpu_upsample_f(const float* src, size_t len, float* dst, size_t samples) {
    for (size_t i = 0; i < samples; ++i) dst[i] = src[(i * len / samples) % len];
}

void refcpu_upsample_hold_c(const float* src, size_t len, float* dst, size_t samples) {
    if (samples == len) { memcpy(dst, src, sizeof(float) * 2 * len); return; }
    const size_t ratio = samples / len;                  /* ipp_upsampler.hpp:35-40 */
    memset(dst, 0, sizeof(float) * 2 * samples);
    for (size_t i = 0; i < len; ++i)
        for (size_t j = 0; j < ratio; ++j) {
            dst[2 * (i * ratio + j)] = src[2 * i];
            dst[2 * (i * ratio + j) + 1] = src[2 * i + 1];
        }
}

size_t refcpu_accumulate_c(const float* src, size_t n, size_t ratio, float* dst) {
    if (ratio == 1) { memcpy(dst, src, sizeof(float) * 2 * n); return n; }
    /* decimator.hpp:57-60: the size is padded to a whole number of blocks
     * whenever there is a remainder */
    size_t q = n / ratio, rem = n % ratio;
    size_t padded = rem ? (q + 1) * ratio : n;
    size_t out_n = padded / ratio;
    for (size_t k = 0; k < out_n; ++k) {
        double sr = 0, si = 0;
        for (size_t j = 0; j < ratio; ++j) {
            size_t i = k * ratio + j;
            if (i < n) { sr += src[2 * i]; si += src[2 * i + 1]; }
        }
        dst[2 * k] = (float)(sr / (double)ratio);
        dst[2 * k + 1] = (float)(si / (double)ratio);
    }
    return out_n;
}

static size_t rc_gcd(size_t a, size_t b) { while (b) { size_t t = a % b; a = b; b = t; } return a; }

size_t refcpu_resample_len(size_t n, size_t new_fs, size_t old_fs) {
    size_t lcm = new_fs / rc_gcd(new_fs, old_fs) * old_fs;
    size_t up = lcm * n / old_fs, ratio = lcm / new_fs;
    return ratio == 1 ? up : (up + ratio - 1) / ratio;
}

size_t refcpu_resample_c(const float* src, size_t n, size_t new_fs, size_t old_fs, float* dst) {
    size_t lcm = new_fs / rc_gcd(new_fs, old_fs) * old_fs;   /* resampler.hpp:41-45 */
    size_t up_n = lcm * n / old_fs;
    size_t ratio = lcm / new_fs;
    float* up = (float*)malloc(sizeof(float) * 2 * up_n);
    refcpu_upsample_hold_c(src, n, up, up_n);
    size_t out_n = refcpu_accumulate_c(up, up_n, ratio, dst);
    free(up);
    return out_n;
}

double refcpu_adjust_sampling_rate(double fs, double target) {
    double new_fs = fs;                                       /* fse.hpp:222-240 */
    size_t lim = (size_t)(2 * fs / target);
    for (size_t i = 1; i < lim; ++i) {
        double cand = fs / (double)i;
        if (fmod(cand / 1e3, 1) != 0.0) continue;
        if (fabs(target - cand) > fabs(target - new_fs)) return new_fs;
        new_fs = cand;
    }
    return fs;
}

/* ============================================================== formats == */
int64_t refcpu_bbp_epoch_bytes(double fs) {             /* signal_parameters.hpp:74-80 */
    double words = fs / 1e3 * 4 / 64 + 1;   /* 2+2 bits a sample, 64-bit words, header */
    int64_t bytes = (int64_t)(words / 2) * 2;
    if (fmod(words, 2.0) != 0.0) bytes += 2;
    return bytes * 8;
}

int refcpu_bbp_verify_headers(const uint8_t* data, int64_t n_epochs, int64_t epoch_bytes) {
    uint64_t last, cur;                                 /* signal_parameters.hpp:36-45 */
    memcpy(&last, data, 8);
    last &= (1u << 24) - 1;
    for (int64_t i = 1; i < n_epochs; ++i) {
        memcpy(&cur, data + i * epoch_bytes, 8);
        cur &= (1u << 24) - 1;
        if (cur - last != 1) return -1;
        last = cur;
    }
    return 0;
}

/* one byte = two samples of signed 2-bit fields, the first declared field in
 * the low bits (BbpPackedSpan.hpp:11-19) */
struct rc_bbp_packed {
    int8_t im_0 : 2;
    int8_t re_0 : 2;
    int8_t im_1 : 2;
    int8_t re_1 : 2;
};

void refcpu_unpack_bbp(const uint8_t* data, int64_t n, int64_t spm, int64_t epoch_bytes, float* out) {
    for (int64_t i = 0; i < n; ++i) {                   /* signal_parameters.hpp:225-236 */
        int64_t e = i / spm, j = i % spm;
        struct rc_bbp_packed pk;
        memcpy(&pk, data + e * epoch_bytes + 8 + j / 2, 1);
        int re = (j & 1) ? pk.re_1 : pk.re_0, im = (j & 1) ? pk.im_1 : pk.im_0;
        out[2 * i] = (float)(re | 1);                   /* BbpPackedSpan.hpp:16-18 */
        out[2 * i + 1] = (float)(im | 1);
    }
}

void refcpu_unpack_nt1065(const uint8_t* data, int64_t n, int offset_bits, float* out) {
    for (int64_t i = 0; i < n; ++i) {                   /* NtlabPackedSpan.hpp:10-21 */
        int sign = (data[i] >> offset_bits) & 1, magnitude = (data[i] >> (offset_bits + 1)) & 1;
        out[2 * i] = (float)((2 * sign - 1) * (1 + 2 * magnitude));
        out[2 * i + 1] = 0.f;
    }
}

/* ================================================================ mixer == */
static double rc_fix_phase(double phase) {               /* mixer.hpp:20-25 */
    phase = fmod(phase, 2 * M_PI);
    if (phase < 0) phase += 2 * M_PI;
    return phase;
}

static void rc_tone_mul(float* x, size_t n, double fs, double f, double phase) {
    const double ph_turns = phase / (2 * M_PI);
    for (size_t i = 0; i < n; ++i) {
        double t = fmod((double)i * f / fs + ph_turns, 1.0);
        double c = cos(2 * M_PI * t), s = sin(2 * M_PI * t);
        double xr = x[2 * i], xi = x[2 * i + 1];
        x[2 * i] = (float)(xr * c - xi * s);
        x[2 * i + 1] = (float)(xr * s + xi * c);
    }
}

void refcpu_mix_c(float* x, size_t n, double fs, double f, double phase) {
    if (f < 0) f = fs + f;                               /* mixer.hpp:46-47 */
    phase = rc_fix_phase(phase);
    rc_tone_mul(x, n, fs, f, phase);
}

void refcpu_mix_inplace_c(float* x, size_t n, double fs, double f, double phase) {
    phase = rc_fix_phase(phase);                         /* mixer.hpp:35-40 */
    if (fabs(f) == 0.0) return;
    if (f < 0) f = fs + f;
    rc_tone_mul(x, n, fs, f, phase);
}

void refcpu_mixer_member_translate_c(float* x, size_t n, double fs, double f, double* phase) {
    refcpu_mix_inplace_c(x, n, fs, f, *phase);           /* mixer.hpp:53-60 */
    double m = fmod(f / 1000.0, 1.0);
    if (m < 0) m += 1;
    *phase += 2 * M_PI * m;
}

/* TableMixer sine table (table_mixer.hpp:31-47): 5/4 of a period in 64 steps
 * per period, FP64 sine rounded once to float. */
void refcpu_table_mixer_table(float* table80) {
    const double step = 5 * M_PI / 80.0 / 2;
    for (int i = 0; i < 80; ++i) table80[i] = (float)(1.0 * sin((double)i * step));
}

/* TableMixer::Process (table_mixer.hpp:78-101) with its NCO (nco.hpp:17-20):
 * the products are rounded separately (no fused multiply-add), as the
 * reference's complex<float> multiply is. */
__attribute__((optimize("fp-contract=off")))
static void rc_table_mul(float* x, size_t n, double fs, double f, double phase) {
    float table[80];
    refcpu_table_mixer_table(table);
    const uint32_t adder = (uint32_t)(uint64_t)(4294967296.0 * f / fs);
    uint32_t holder = (uint32_t)(uint64_t)(4294967296.0 * phase / (2 * M_PI));
    for (size_t i = 0; i < n; ++i) {
        const uint32_t idx = holder >> 26;             /* table_mixer.hpp:53 */
        holder += adder;                               /* 32-bit accumulator wraps */
        const float c = table[idx + 16], s = table[idx];   /* cos at +1/4 period */
        const float a = x[2 * i], b = x[2 * i + 1];
        const float ac = a * c, bs = b * s, as = a * s, bc = b * c;
        x[2 * i] = ac - bs;
        x[2 * i + 1] = as + bc;
    }
}

void refcpu_table_mix_c(float* x, size_t n, double fs, double f, double phase) {
    phase = rc_fix_phase(phase);                         /* mixer.hpp:35-40 */
    if (fabs(f) == 0.0) return;
    if (f < 0) f = fs + f;
    rc_table_mul(x, n, fs, f, phase);
}

void refcpu_table_mixer_member_c(float* x, size_t n, double fs, double f, double* phase) {
    refcpu_table_mix_c(x, n, fs, f, *phase);             /* mixer.hpp:53-60 */
    double m = fmod(f / 1000.0, 1.0);
    if (m < 0) m += 1;
    *phase += 2 * M_PI * m;
}

/* ======================================================= matched filter == */
int refcpu_mf_prepare_f(const float* code, size_t n, float* spec) {
    for (size_t i = 0; i < n; ++i) { spec[2 * i] = code[i]; spec[2 * i + 1] = 0.f; }
    if (refcpu_fft_f(spec, n, 0)) return -1;
    for (size_t i = 0; i < n; ++i) spec[2 * i + 1] = -spec[2 * i + 1];
    return 0;
}

int refcpu_mf_process_f(float* x, const float* spec, size_t n) {
    if (refcpu_fft_f(x, n, 0)) return -1;
    for (size_t i = 0; i < n; ++i) {
        float ar = x[2 * i], ai = x[2 * i + 1], br = spec[2 * i], bi = spec[2 * i + 1];
        x[2 * i] = ar * br - ai * bi;
        x[2 * i + 1] = ar * bi + ai * br;
    }
    return refcpu_fft_f(x, n, 1);
}

/* ================================================================= math == */
void refcpu_abs_c(const float* x, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) out[i] = (float)hypot((double)x[2 * i], (double)x[2 * i + 1]);
}

int refcpu_reshape_sum_f(const float* x, size_t n, size_t block, float* out) {
    if (block == 0 || n % block) return -1;              /* reshape_and_sum.hpp:16-17 */
    memcpy(out, x, sizeof(float) * block);
    for (size_t b = 1; b < n / block; ++b)
        for (size_t j = 0; j < block; ++j) out[j] += x[b * block + j];
    return 0;
}

void refcpu_max_index_f(const float* x, size_t n, float* value, int64_t* index) {
    size_t best = 0;
    for (size_t i = 1; i < n; ++i)
        if (x[i] > x[best]) best = i;
    *value = n ? x[best] : 0.f;
    *index = (int64_t)best;
}

void refcpu_mean_std_f(const float* x, size_t n, double* mean, double* sigma) {
    if (n < 2) { *mean = 0; *sigma = 0; return; }        /* mean_stddev.hpp:33-34 */
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += x[i];
    double m = s / (double)n, q = 0;
    for (size_t i = 0; i < n; ++i) q += ((double)x[i] - m) * ((double)x[i] - m);
    *mean = m;
    *sigma = sqrt(q / (double)(n - 1));
}

/* ========================================================== acquisition == */
typedef struct {
    const refcpu_acq_params* p;
    const float* signal;
    int precision;
    refcpu_acq_result* out;
    float* surfaces;
    float* peaks;
    atomic_long next;
    atomic_int err;
} rc_acq_job;

#define RC_ACQ_BODY(REAL, SUF)                                                                     \
    static int rc_acq_prn##SUF(rc_acq_job* job, int64_t prn, REAL* spec, REAL* buf, REAL* work,    \
                               REAL* tmp_c, double* surf, double* best_surf) {                      \
        const refcpu_acq_params* p = job->p;                                                       \
        const size_t N = (size_t)p->n_fft, L = (size_t)p->surface_len;                             \
        const size_t R = N / L;                                                                    \
        rc_plan##SUF* plan = rc_get_plan##SUF(N);                                  \
        const float* rep = p->replicas + (size_t)prn * N;                                          \
        for (size_t i = 0; i < N; ++i) { spec[2 * i] = (REAL)rep[i]; spec[2 * i + 1] = 0; }        \
        rc_exec##SUF(plan, spec, work, 0);                                                 \
        for (size_t i = 0; i < N; ++i) spec[2 * i + 1] = -spec[2 * i + 1];                         \
        refcpu_acq_result best;                                                                    \
        memset(&best, 0, sizeof(best));                                                            \
        best.sigma = 1.0; /* AcquisitionResult defaults, acquisition_result.hpp:13-18 */           \
        best.doppler_bin = -1;                                                                     \
        double runner_dop = 0;                                                                     \
        for (int64_t di = 0; di < p->n_doppler; ++di) {                                            \
            const double d = p->doppler[di];                                                       \
            double f = -d;                                                                         \
            if (f < 0) f = p->fs_acq + f;                                                          \
            for (size_t j = 0; j < L; ++j) surf[j] = 0;                                            \
            for (int64_t k = 0; k < p->n_blocks; ++k) {                                            \
                const float* x = job->signal + 2 * (size_t)k * N;                                  \
                /* tone exp(j 2 pi f n / fs): FP64 rotator re-seeded exactly every 256 */      \
                const double wr = cos(2 * M_PI * f / p->fs_acq), wi = sin(2 * M_PI * f / p->fs_acq); \
                double c = 1.0, s = 0.0;                                                           \
                for (size_t n = 0; n < N; ++n) {                                                   \
                    if ((n & 255) == 0) {                                                          \
                        double t = fmod((double)n * f / p->fs_acq, 1.0);                           \
                        c = cos(2 * M_PI * t); s = sin(2 * M_PI * t);                              \
                    }                                                                              \
                    REAL cr = (REAL)c, ci = (REAL)s;                                               \
                    REAL xr = (REAL)x[2 * n], xi = (REAL)x[2 * n + 1];                             \
                    buf[2 * n] = xr * cr - xi * ci;                                                \
                    buf[2 * n + 1] = xr * ci + xi * cr;                                            \
                    const double c2 = c * wr - s * wi;                                             \
                    s = c * wi + s * wr;                                                           \
                    c = c2;                                                                        \
                }                                                                                  \
                rc_exec##SUF(plan, buf, work, 0);                                          \
                for (size_t n = 0; n < N; ++n) {                                                   \
                    REAL ar = buf[2 * n], ai = buf[2 * n + 1], br = spec[2 * n], bi = spec[2 * n + 1]; \
                    buf[2 * n] = ar * br - ai * bi;                                                \
                    buf[2 * n + 1] = ar * bi + ai * br;                                            \
                }                                                                                  \
                rc_exec##SUF(plan, buf, work, 1);                                          \
                if (!p->reshape || p->coherent) {                                                  \
                    for (size_t j = 0; j < L; ++j) {                                               \
                        REAL acc = 0;                                                              \
                        for (size_t r = 0; r < R; ++r) {                                           \
                            REAL a = (REAL)sqrt((double)buf[2 * (j + r * L)] * buf[2 * (j + r * L)] + \
                                                (double)buf[2 * (j + r * L) + 1] * buf[2 * (j + r * L) + 1]); \
                            acc = r ? acc + a : a;                                                 \
                        }                                                                          \
                        surf[j] += (double)acc;                                                    \
                    }                                                                              \
                } else {                                                                           \
                    for (size_t j = 0; j < L; ++j) {                                               \
                        REAL sr = 0, si = 0;                                                       \
                        for (size_t r = 0; r < R; ++r) {                                           \
                            sr += buf[2 * (j + r * L)];                                            \
                            si += buf[2 * (j + r * L) + 1];                                        \
                        }                                                                          \
                        tmp_c[j] = (REAL)sqrt((double)sr * sr + (double)si * si);                  \
                        surf[j] += (double)tmp_c[j];                                               \
                    }                                                                              \
                }                                                                                  \
            }                                                                                      \
            if (precision_is_f32##SUF)                                                             \
                for (size_t j = 0; j < L; ++j) surf[j] = (double)(float)surf[j];                   \
            size_t mi = 0;                                                                         \
            for (size_t j = 1; j < L; ++j)                                                         \
                if (surf[j] > surf[mi]) mi = j;                                                    \
            double sum = 0;                                                                        \
            for (size_t j = 0; j < L; ++j) sum += surf[j];                                         \
            double mean = sum / (double)L, q = 0;                                                  \
            for (size_t j = 0; j < L; ++j) q += (surf[j] - mean) * (surf[j] - mean);               \
            double sd = L > 1 ? sqrt(q / (double)(L - 1)) : 0;                                     \
            if (L < 2) mean = 0;                                                                   \
            if (job->surfaces) {                                                                   \
                float* o = job->surfaces + ((size_t)prn * p->n_doppler + di) * L;                  \
                for (size_t j = 0; j < L; ++j) o[j] = (float)surf[j];                              \
            }                                                                                      \
            double level = surf[mi];                                                               \
            if (best.level < level) { /* strict, fse.hpp:265 */                                    \
                if (best.doppler_bin >= 0 && best.level > runner_dop) runner_dop = best.level;     \
                best.level = level;                                                                \
                best.mean = mean;                                                                  \
                best.stddev = sd;                                                                  \
                best.sigma = sd + mean;                                                            \
                best.index = (int64_t)mi;                                                          \
                best.code_offset = p->fs_in / p->fs_acq * (double)mi;                              \
                best.doppler = d + p->intermediate_frequency;                                      \
                best.doppler_bin = (int32_t)di;                                                    \
                double r2 = 0;                                                                     \
                for (size_t j = 0; j < L; ++j)                                                     \
                    if (j != mi && surf[j] > r2) r2 = surf[j];                                     \
                best.runner_up_code = r2;                                                          \
                for (size_t j = 0; j < L; ++j) best_surf[j] = surf[j];                             \
            } else if (level > runner_dop) {                                                       \
                runner_dop = level;                                                                \
            }                                                                                      \
        }                                                                                          \
        best.runner_up_doppler = runner_dop;                                                       \
        best.snr = best.level / best.sigma;                                                        \
        best.detected = (!p->reshape || best.snr > p->threshold) ? 1 : 0;                         \
        job->out[prn] = best;                                                                      \
        if (job->peaks)                                                                            \
            for (size_t j = 0; j < L; ++j)                                                         \
                job->peaks[(size_t)prn * L + j] = best.doppler_bin >= 0 ? (float)best_surf[j] : 0.f; \
        return 0;                                                                                  \
    }

enum { precision_is_f32_d = 0, precision_is_f32_f = 1 };
RC_ACQ_BODY(double, _d)
RC_ACQ_BODY(float, _f)

static void* rc_acq_worker(void* arg) {
    rc_acq_job* job = (rc_acq_job*)arg;
    const refcpu_acq_params* p = job->p;
    const size_t N = (size_t)p->n_fft, L = (size_t)p->surface_len;
    const size_t esz = job->precision == 64 ? sizeof(double) : sizeof(float);
    char* mem = (char*)malloc(esz * 2 * N * 3 + esz * 2 * L + sizeof(double) * 2 * L);
    if (!mem) { atomic_store(&job->err, 1); return NULL; }
    void* spec = mem;
    void* buf = mem + esz * 2 * N;
    void* work = mem + esz * 4 * N;
    void* tmp_c = mem + esz * 6 * N;
    double* surf = (double*)(mem + esz * 6 * N + esz * 2 * L);
    double* best_surf = surf + L;
    for (;;) {
        long prn = atomic_fetch_add(&job->next, 1);
        if (prn >= p->n_prn) break;
        if (job->precision == 64)
            rc_acq_prn_d(job, prn, (double*)spec, (double*)buf, (double*)work, (double*)tmp_c, surf, best_surf);
        else
            rc_acq_prn_f(job, prn, (float*)spec, (float*)buf, (float*)work, (float*)tmp_c, surf, best_surf);
    }
    free(mem);
    return NULL;
}

int refcpu_acquire(const refcpu_acq_params* p, const float* signal, int precision, int threads,
                   refcpu_acq_result* out, float* surfaces, float* peaks) {
    if (!p || !signal || !out || p->n_fft <= 0 || p->surface_len <= 0 || p->n_blocks <= 0) return -1;
    if (p->n_fft % p->surface_len) return -2;            /* reshape_and_sum.hpp:16-17 */
    if (!p->reshape && p->surface_len != p->n_fft) return -3;
    if (precision != 32 && precision != 64) return -4;
    rc_acq_job job;
    job.p = p; job.signal = signal; job.precision = precision; job.out = out;
    job.surfaces = surfaces; job.peaks = peaks;
    atomic_init(&job.next, 0);
    atomic_init(&job.err, 0);
    /* build the plan once before fanning out */
    if (precision == 64) rc_get_plan_d((size_t)p->n_fft); else rc_get_plan_f((size_t)p->n_fft);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, rc_acq_worker, &job);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    return atomic_load(&job.err) ? -5 : 0;
}

/* =========================================================== correlator == */
void refcpu_correlate(const float* signal, const float* code, size_t n, double* out) {
    double r = 0, i = 0;
    for (size_t k = 0; k < n; ++k) { r += (double)signal[2 * k] * code[k]; i += (double)signal[2 * k + 1] * code[k]; }
    out[0] = r; out[1] = i;
}

static void rc_add_with_phase(const double* l0, const double* r0, double rel, double* out) {
    double lhs[2] = {l0[0], l0[1]}, rhs[2] = {r0[0], r0[1]};
    if (rel < 0.5) {                                     /* tracking_parameters.hpp:89-98 */
        double t0 = lhs[0], t1 = lhs[1];
        lhs[0] = rhs[0]; lhs[1] = rhs[1]; rhs[0] = t0; rhs[1] = t1;
    }
    double wl = lhs[0] / rel, wr = rhs[0] / (1.0 - rel);
    if (fabs(wl + wr) > fabs(wl)) { out[0] = lhs[0] + rhs[0]; out[1] = lhs[1] + rhs[1]; }
    else { out[0] = lhs[0] - rhs[0]; out[1] = lhs[1] - rhs[1]; }
}

void refcpu_correlate_split(const float* signal, size_t spm, const float* code3, double code_period_ms,
                            double phase, double* out) {
    const double cps = (double)spm * code_period_ms;     /* tracking_parameters.hpp:471-493 */
    while (phase < 0) phase += cps;
    if (phase > cps) phase = fmod(phase, cps);
    size_t first_len = (size_t)ceil(phase) % spm;
    size_t second_len = spm - first_len;
    size_t first_phase = (size_t)ceil(2 * cps - phase);
    size_t second_phase = first_phase + first_len;
    double a[2], b[2];
    refcpu_correlate(signal, code3 + first_phase, first_len, a);
    refcpu_correlate(signal + 2 * first_len, code3 + second_phase, second_len, b);
    rc_add_with_phase(a, b, fmod(phase, (double)spm) / (double)spm, out);
}

void refcpu_epl(float* x, size_t spm, double fs, double carrier_frequency, double carrier_phase,
                const float* code3, double code_period_ms, double code_phase, double samples_per_chip,
                const double* taps_chips, int n_taps, double* out) {
    refcpu_mix_inplace_c(x, spm, fs, -carrier_frequency, -carrier_phase);   /* tracker.hpp:121 */
    for (int t = 0; t < n_taps; ++t)                                          /* tracker.hpp:126-136 */
        refcpu_correlate_split(x, spm, code3, code_period_ms, code_phase + samples_per_chip * taps_chips[t],
                               out + 2 * t);
}

/* loop gains (tracking_parameters.hpp:72-87) */
static const double RC_PLL_W = 25.0 / 0.53;
#define RC_K1_PLL (0.001 * (RC_PLL_W * RC_PLL_W) + 1.414 * RC_PLL_W)
#define RC_K2_PLL (1.414 * RC_PLL_W)
#define RC_K3_PLL (0.001 * (250.0 / 0.25))
static const double RC_DLL_WN = 0.7 * 8.0 * 1.0 / (4.0 * 1.0 * 1.0 + 1.0);
#define RC_DLL_TAU1 (0.1 / (RC_DLL_WN * RC_DLL_WN))
#define RC_DLL_TAU2 (2.0 * 1.0 / RC_DLL_WN)

static void rc_pll(refcpu_track_state* s, float pi_, float pq) {   /* tracking_parameters.hpp:495-511 */
    double e = (pi_ != 0.0f) ? atan((double)(pq / pi_)) / (M_PI * 2.0) : 0.0;
    float qi = s->previous_prompt[0], qq = s->previous_prompt[1];
    float cross = pi_ * qq - qi * pq;
    float dot = fabsf(pi_ * qi + qq * pq);
    double ferr = (double)atan2f(cross, dot) / M_PI;
    s->carrier_frequency += RC_K1_PLL * e - RC_K2_PLL * s->carrier_phase_error - RC_K3_PLL * ferr;
    s->carrier_phase_error = e;
    s->carrier_phase += e * M_PI / 2;
    s->previous_prompt[0] = pi_;
    s->previous_prompt[1] = pq;
}

static void rc_dll(refcpu_track_state* s, const float* e, const float* l, double fs) {  /* :513-528 */
    float ea = hypotf(e[0], e[1]), la = hypotf(l[0], l[1]);
    double err = 0.0;
    if (ea + la != 0.0f) err = (ea - la) / (ea + la);
    s->code_nco += (RC_DLL_TAU2 / RC_DLL_TAU1 * (err - s->code_error) + 0.001 / RC_DLL_TAU1 * err) *
                   (s->base_code_frequency / 1.023e6);
    s->code_error = err;
    s->code_frequency = s->base_code_frequency - s->code_nco;
    s->code_phase -= fs / 1000 / 2 * (s->code_frequency / s->base_code_frequency - 1);
}

void refcpu_track(const float* x, int64_t spm, int64_t n_blocks, double fs, const float* code3, double spacing,
                  refcpu_track_state* st, float* epl, double* trace) {
    float* blk = (float*)malloc(sizeof(float) * 2 * (size_t)spm);
    const double taps[3] = {spacing, 0.0, -spacing};
    for (int64_t b = 0; b < n_blocks; ++b) {            /* tracker.hpp:156-177 */
        memcpy(blk, x + 2 * (size_t)b * spm, sizeof(float) * 2 * (size_t)spm);
        double cp = st->code_phase - fs / 1e3 * (double)(st->epochs % st->code_period_ms);
        double out[6];
        refcpu_epl(blk, (size_t)spm, fs, st->carrier_frequency, st->carrier_phase, code3,
                   (double)st->code_period_ms, cp, fs / st->base_code_frequency, taps, 3, out);
        st->carrier_phase += 2 * M_PI * st->carrier_frequency / 1000.0;   /* UpdatePhase */
        float v[6];
        for (int i = 0; i < 6; ++i) v[i] = (float)out[i];
        memcpy(epl + 6 * (size_t)b, v, sizeof v);
        rc_pll(st, v[2], v[3]);
        rc_dll(st, v, v + 4, fs);
        st->epochs += 1;
        if (trace) {
            trace[4 * b] = -st->carrier_phase / (2 * M_PI);
            trace[4 * b + 1] = st->carrier_frequency;
            trace[4 * b + 2] = st->code_phase;
            trace[4 * b + 3] = st->code_frequency;
        }
    }
    free(blk);
}

typedef struct {
    const float* x; int64_t spm, n_blocks; double fs;
    const refcpu_channel* ch; int n_ch; const int8_t* codes; int64_t code_len;
    const double* taps; int n_taps; double* out;
    atomic_int next;
} rc_bank_job;

static void* rc_bank_worker(void* arg) {
    rc_bank_job* j = (rc_bank_job*)arg;
    for (;;) {
        int c = atomic_fetch_add(&j->next, 1);
        if (c >= j->n_ch) break;
        const refcpu_channel ch = j->ch[c];
        const int8_t* code = j->codes + (size_t)c * j->code_len;
        const double L = (double)j->code_len;
        for (int64_t b = 0; b < j->n_blocks; ++b) {
            const double N = (double)j->spm;
            double theta = ch.carrier_phase + ch.carrier_frequency * (double)b * N / j->fs;
            double tau = ch.code_phase + ch.code_rate * (double)b * N / j->fs;
            theta -= floor(theta);
            tau = fmod(tau, L);
            double acc[64][2];
            for (int t = 0; t < j->n_taps; ++t) acc[t][0] = acc[t][1] = 0;
            const float* xb = j->x + 2 * (size_t)b * j->spm;
            for (int64_t n = 0; n < j->spm; ++n) {
                double ph = theta + ch.carrier_frequency * (double)n / j->fs;
                ph -= floor(ph);
                double c0 = cos(2 * M_PI * ph), s0 = sin(2 * M_PI * ph);
                double xr = xb[2 * n], xi = xb[2 * n + 1];
                double sr = xr * c0 + xi * s0, si = xi * c0 - xr * s0;
                double u = tau + ch.code_rate * (double)n / j->fs;
                for (int t = 0; t < j->n_taps; ++t) {
                    double v = floor(u + j->taps[t]);
                    int64_t idx = (int64_t)fmod(v, L);
                    if (idx < 0) idx += j->code_len;
                    acc[t][0] += sr * code[idx];
                    acc[t][1] += si * code[idx];
                }
            }
            double* o = j->out + 2 * (((size_t)b * j->n_ch + c) * j->n_taps);
            for (int t = 0; t < j->n_taps; ++t) { o[2 * t] = acc[t][0]; o[2 * t + 1] = acc[t][1]; }
        }
    }
    return NULL;
}

void refcpu_bank_nco(const float* x, int64_t spm, int64_t n_blocks, double fs, const refcpu_channel* ch, int n_ch,
                     const int8_t* codes, int64_t code_len, const double* taps_chips, int n_taps, int threads,
                     double* out) {
    if (n_taps > 64) return;
    rc_bank_job j = {x, spm, n_blocks, fs, ch, n_ch, codes, code_len, taps_chips, n_taps, out, 0};
    atomic_init(&j.next, 0);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, rc_bank_worker, &j);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}
