/*
 * Deterministic GNSS baseband signal synthesiser used for the acquisition and
 * tracking workloads (synthetic data; no recording is needed).
 *
 *   s[n] = sum_i A_i b_i(t) c_i(floor(t r_i - tau_i) mod L_i) exp(j 2 pi (f_i t + theta_i)) + w[n]
 *
 * t = n / fs, r_i = chip_rate_i (1 + f_i / carrier_hz_i) (code Doppler),
 * A_i = sigma sqrt(2 10^(CN0_i/10) / fs), w complex AWGN with per-component
 * sigma (Box-Muller in FP64 over xoshiro256** seeded through splitmix64),
 * b_i = +-1 navigation bits of period nav_period_s starting at nav_offset_s.
 * Output is rounded to nearest and clipped to int8 / int16, or kept as float.
 *
 * Randomness is counter based (one generator per 4096-sample chunk), so the
 * result does not depend on the thread count.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    const int8_t* code;     /* +-1 chips (already BOC expanded when BOC) */
    int64_t code_len;
    double chip_rate;       /* nominal chips/s of `code`                 */
    double carrier_hz;      /* for code Doppler: r = chip_rate(1+f/carrier) */
    double code_phase;      /* tau, chips of delay at t = 0             */
    double doppler;         /* Hz                                        */
    double carrier_phase;   /* cycles                                    */
    double cn0_dbhz;
    double nav_period_s;    /* 0: no navigation bits                     */
    double nav_offset_s;
    uint64_t nav_seed;
} ugsdr_siggen_sat;

static inline uint64_t sg_splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } sg_xo;

static inline uint64_t sg_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t sg_next(sg_xo* r) {              /* xoshiro256** */
    const uint64_t res = sg_rotl(r->s[1] * 5, 7) * 9;
    const uint64_t t = r->s[1] << 17;
    r->s[2] ^= r->s[0]; r->s[3] ^= r->s[1]; r->s[1] ^= r->s[2]; r->s[0] ^= r->s[3];
    r->s[2] ^= t;
    r->s[3] = sg_rotl(r->s[3], 45);
    return res;
}

static inline void sg_seed(sg_xo* r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = sg_splitmix(&x);
}

static inline double sg_uniform(sg_xo* r) {             /* (0,1) */
    return ((double)(sg_next(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* +-1 code of len chips from a seed */
void ugsdr_siggen_random_code(uint64_t seed, int64_t len, int8_t* out) {
    sg_xo r;
    sg_seed(&r, seed ^ 0xC0DEC0DEull);
    for (int64_t i = 0; i < len; ++i) out[i] = (sg_next(&r) >> 63) ? 1 : -1;
}

static inline int sg_nav_bit(uint64_t seed, int64_t k) {
    uint64_t x = seed ^ ((uint64_t)k * 0xD1B54A32D192ED03ull);
    return (sg_splitmix(&x) >> 63) ? 1 : -1;
}

typedef struct {
    const ugsdr_siggen_sat* sats;
    int n_sats;
    double fs;
    int64_t n;
    double sigma;
    uint64_t seed;
    int fmt;
    void* out;
    atomic_long next;
} sg_job;

#define SG_CHUNK 4096

static void* sg_worker(void* arg) {
    sg_job* j = (sg_job*)arg;
    double* re = (double*)malloc(sizeof(double) * 2 * SG_CHUNK);
    double* im = re + SG_CHUNK;
    const int64_t n_chunks = (j->n + SG_CHUNK - 1) / SG_CHUNK;
    for (;;) {
        int64_t ck = atomic_fetch_add(&j->next, 1);
        if (ck >= n_chunks) break;
        const int64_t n0 = ck * SG_CHUNK;
        const int64_t m = (j->n - n0) < SG_CHUNK ? (j->n - n0) : SG_CHUNK;
        /* noise */
        sg_xo r;
        sg_seed(&r, j->seed * 0x9E3779B97F4A7C15ull + (uint64_t)ck);
        for (int64_t i = 0; i < m; ++i) {
            double u1 = sg_uniform(&r), u2 = sg_uniform(&r);
            double rad = sqrt(-2.0 * log(u1)) * j->sigma;
            re[i] = rad * cos(2 * M_PI * u2);
            im[i] = rad * sin(2 * M_PI * u2);
        }
        /* satellites */
        for (int s = 0; s < j->n_sats; ++s) {
            const ugsdr_siggen_sat* st = &j->sats[s];
            const double amp = j->sigma * sqrt(2.0 * pow(10.0, st->cn0_dbhz / 10.0) / j->fs);
            const double rate = st->chip_rate * (1.0 + st->doppler / st->carrier_hz);
            const double L = (double)st->code_len;
            /* exact carrier phasor at chunk start, then a rotator */
            double ph = st->carrier_phase + st->doppler * (double)n0 / j->fs;
            ph -= floor(ph);
            double cr = cos(2 * M_PI * ph), ci = sin(2 * M_PI * ph);
            const double dr = cos(2 * M_PI * st->doppler / j->fs), di = sin(2 * M_PI * st->doppler / j->fs);
            for (int64_t i = 0; i < m; ++i) {
                const double t = (double)(n0 + i) / j->fs;
                double u = t * rate - st->code_phase;
                double chip = floor(u);
                int64_t idx = (int64_t)fmod(chip, L);
                if (idx < 0) idx += st->code_len;
                double v = amp * (double)st->code[idx];
                if (st->nav_period_s > 0) {
                    int64_t k = (int64_t)floor((t - st->nav_offset_s) / st->nav_period_s);
                    v *= (double)sg_nav_bit(st->nav_seed, k);
                }
                re[i] += v * cr;
                im[i] += v * ci;
                const double t0 = cr * dr - ci * di;
                ci = cr * di + ci * dr;
                cr = t0;
            }
        }
        /* quantise */
        if (j->fmt == 8) {
            int8_t* o = (int8_t*)j->out + 2 * n0;
            for (int64_t i = 0; i < m; ++i) {
                double a = nearbyint(re[i]), b = nearbyint(im[i]);
                a = a > 127 ? 127 : (a < -127 ? -127 : a);
                b = b > 127 ? 127 : (b < -127 ? -127 : b);
                o[2 * i] = (int8_t)a; o[2 * i + 1] = (int8_t)b;
            }
        } else if (j->fmt == 16) {
            int16_t* o = (int16_t*)j->out + 2 * n0;
            for (int64_t i = 0; i < m; ++i) {
                double a = nearbyint(re[i]), b = nearbyint(im[i]);
                a = a > 32767 ? 32767 : (a < -32767 ? -32767 : a);
                b = b > 32767 ? 32767 : (b < -32767 ? -32767 : b);
                o[2 * i] = (int16_t)a; o[2 * i + 1] = (int16_t)b;
            }
        } else {
            float* o = (float*)j->out + 2 * n0;
            for (int64_t i = 0; i < m; ++i) { o[2 * i] = (float)re[i]; o[2 * i + 1] = (float)im[i]; }
        }
    }
    free(re);
    return NULL;
}

/* fmt: 8 (int8 IQ), 16 (int16 IQ) or 32 (float IQ).  out: 2*n elements. */
int ugsdr_siggen_generate(const ugsdr_siggen_sat* sats, int n_sats, double fs, int64_t n, double sigma,
                          uint64_t seed, int fmt, void* out, int threads) {
    if (n < 0 || (fmt != 8 && fmt != 16 && fmt != 32) || (!out && n) || fs <= 0) return -1;
    sg_job j;
    j.sats = sats; j.n_sats = n_sats; j.fs = fs; j.n = n; j.sigma = sigma; j.seed = seed;
    j.fmt = fmt; j.out = out;
    atomic_init(&j.next, 0);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, sg_worker, &j);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}
