// Context lifecycle, error reporting, host PRN code generators and the
// device-resident replica-spectrum cache.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ctx.hpp"
#include "internal.hpp"

namespace ugsdr_hip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

hipError_t set_lds_limit(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, size_t> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    size_t& have = done[{kernel, dev}];
    if (have >= bytes) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) have = bytes;
    return e;
}

}  // namespace ugsdr_hip

using namespace ugsdr_hip;

int64_t ugsdr_hip_ctx::device_bytes() const {
    int64_t b = 0;
    for (const DevBuf* d : {&iq, &sig, &classes, &scratch, &rows, &results, &surf, &tables, &aux, &aux2})
        b += (int64_t)d->bytes;
    b += slot.bytes();
    b += (int64_t)track.buf.bytes;
    for (const auto& kv : replicas)
        b += kv.second.n_fft * (kv.second.precision == UGSDR_PREC_FP32 ? 8 : 4) +
             (kv.second.spec_short ? kv.second.period * 8 : 0) + kv.second.code_len * 8;
    for (const auto& kv : boxcar) b += (int64_t)kv.second.bytes;
    return b;
}

extern "C" {

const char* ugsdr_hip_last_error(void) { return g_last_error.c_str(); }

const char* ugsdr_hip_version(void) { return "ugsdr_hip 0.1 gfx950"; }

int ugsdr_hip_ctx_create(int device, ugsdr_hip_ctx** out) {
    if (!out) {
        set_error("null output pointer");
        return UGSDR_ERR_ARG;
    }
    *out = nullptr;
    int count = 0;
    UGSDR_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        set_error("no such device");
        return UGSDR_ERR_ARG;
    }
    UGSDR_HIP_TRY(hipSetDevice(device));
    if (int rc = noalloc_scratch_init(device)) return rc;
    auto* ctx = new ugsdr_hip_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return UGSDR_ERR_HIP;
    }
    *out = ctx;
    return UGSDR_OK;
}

int ugsdr_hip_ctx_destroy(ugsdr_hip_ctx* ctx) {
    if (!ctx) return UGSDR_OK;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->replicas) {
            if (kv.second.spec) (void)hipFree(kv.second.spec);
            if (kv.second.spec_short) (void)hipFree(kv.second.spec_short);
            if (kv.second.code_spec) (void)hipFree(kv.second.code_spec);
        }
        for (auto& kv : ctx->boxcar) kv.second.release();
        for (auto& e : ctx->ev)
            if (e) (void)hipEventDestroy(e);
        ctx->replicas.clear();
        for (DevBuf* d : {&ctx->iq, &ctx->sig, &ctx->classes, &ctx->scratch, &ctx->rows, &ctx->results, &ctx->surf,
                          &ctx->tables, &ctx->aux, &ctx->aux2})
            d->release();
        ctx->graph.drop();
        ctx->slot.release_all();
        ctx->track.buf.release();
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return UGSDR_OK;
}

int ugsdr_hip_ctx_stream(ugsdr_hip_ctx* ctx, void** stream) {
    if (!ctx || !stream) {
        set_error("null argument");
        return UGSDR_ERR_ARG;
    }
    *stream = ctx->stream;
    return UGSDR_OK;
}

int ugsdr_hip_ctx_set_profiling(ugsdr_hip_ctx* ctx, int enable) {
    if (!ctx) {
        set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (enable && !ctx->ev[0])
        for (auto& e : ctx->ev) UGSDR_HIP_TRY(hipEventCreate(&e));
    ctx->profile = enable != 0;
    return UGSDR_OK;
}

int ugsdr_hip_ctx_last_times(ugsdr_hip_ctx* ctx, float* ms) {
    if (!ctx || !ms) {
        set_error("null argument");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int i = 0; i < 4; ++i) ms[i] = ctx->last_ms[i];
    return UGSDR_OK;
}

int64_t ugsdr_hip_ctx_device_bytes(ugsdr_hip_ctx* ctx) {
    if (!ctx) return 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return ctx->device_bytes();
}

// ------------------------------------------------------------------ codes
// Fibonacci LFSR with the reference's conventions: the output is bit `outpin`,
// the feedback is the parity of (value & poly) shifted in at bit 0, and an
// optional short cycle reloads `reset_value` after `reset_period` shifts.
struct Lfsr27 {
    uint32_t value, poly, outpin, reset_period = 0, reset_value = 0, count = 0;
    static constexpr uint32_t kMask = (1u << 27) - 1;
    Lfsr27(uint32_t v, uint32_t p, uint32_t o) : value(v & kMask), poly(p & kMask), outpin(o) {}
    int out() const { return (int)((value >> outpin) & 1u); }
    void shift() {
        ++count;
        if (reset_period && count == reset_period) {
            count = 0;
            value = reset_value;
            return;
        }
        const uint32_t fb = (uint32_t)__builtin_parity(value & poly);
        value = ((value << 1) | fb) & kMask;
    }
    // 2^(bit length of poly) - 1
    uint32_t period() const {
        uint32_t bits = 0, p = poly;
        while (p) { ++bits; p >>= 1; }
        return (1u << bits) - 1;
    }
};

int ugsdr_code_gps_l1ca(int sv, int8_t* out) {
    // G2 delays per ICD-GPS-200 (same table as src/prn_codes/GpsL1Ca.hpp:10-15)
    static const short delays[37] = {5,   6,   7,   8,   17,  18,  139, 140, 141, 251, 252, 254, 255,
                                     256, 257, 258, 469, 470, 471, 472, 473, 474, 509, 512, 513, 514,
                                     515, 516, 859, 860, 861, 862, 863, 950, 947, 948, 950};
    if (!out) { set_error("null output"); return UGSDR_ERR_ARG; }
    if (sv < 0 || sv >= 37) {
        set_error("SV number for the GPS C/A code exceed available values");
        return UGSDR_ERR_ARG;
    }
    Lfsr27 g1(0x3FF, 0x204, 9), g2(0x3FF, 0x3A6, 9);
    for (int i = 0; i < 1023 - delays[sv]; ++i) g2.shift();
    const uint32_t n = g1.period();
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = (int8_t)(1 - 2 * (g1.out() ^ g2.out()));
        g1.shift();
        g2.shift();
    }
    return UGSDR_OK;
}

static int l5_code(int advance, int8_t* out) {
    Lfsr27 xa(0xFFFF, 0x1B00, 12), xb(0xFFFF, 0x18ED, 12);
    xa.reset_period = 8190;
    xa.reset_value = 0xFFFF;
    for (int i = 0; i < advance; ++i) xb.shift();
    for (int i = 0; i < 10230; ++i) {
        out[i] = (int8_t)(1 - 2 * (xa.out() ^ xb.out()));
        xa.shift();
        xb.shift();
    }
    return UGSDR_OK;
}

int ugsdr_code_gps_l5i(int sv, int8_t* out) {
    // XB advances per IS-GPS-705 (same table as src/prn_codes/GpsL5I.hpp:10-15)
    static const short adv[32] = {266,  365,  804,  1138, 1509, 1559, 1756, 2084, 2170, 2303, 2527,
                                  2687, 2930, 3471, 3940, 4132, 4332, 4924, 5343, 5443, 5641, 5816,
                                  5898, 5918, 5955, 6243, 6345, 6477, 6518, 6875, 7168, 7187};
    if (!out) { set_error("null output"); return UGSDR_ERR_ARG; }
    if (sv < 0 || sv >= 32) {
        set_error("SV number for the GPS L5I code exceed available values");
        return UGSDR_ERR_ARG;
    }
    return l5_code(adv[sv], out);
}

int ugsdr_code_gps_l5q(int sv, int8_t* out) {
    // XB advances per IS-GPS-705 (same table as src/prn_codes/GpsL5Q.hpp:10-15)
    static const short adv[32] = {1701, 323,  5292, 2020, 5429, 7136, 1041, 5947, 4315, 148,  535,
                                  1939, 5206, 5910, 3595, 5135, 6082, 6990, 3546, 1523, 4548, 4484,
                                  1893, 3961, 7106, 5299, 4660, 276,  4389, 3783, 1591, 1601};
    if (!out) { set_error("null output"); return UGSDR_ERR_ARG; }
    if (sv < 0 || sv >= 32) {
        set_error("SV number for the GPS L5Q code exceed available values");
        return UGSDR_ERR_ARG;
    }
    return l5_code(adv[sv], out);
}

int ugsdr_code_glonass_of(int8_t* out) {
    if (!out) { set_error("null output"); return UGSDR_ERR_ARG; }
    Lfsr27 g(0x1FF, 0x110, 6);
    const uint32_t n = g.period();
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = (int8_t)(1 - 2 * g.out());
        g.shift();
    }
    return UGSDR_OK;
}

int ugsdr_code_boc(const int8_t* chips, int64_t n, int8_t* out) {
    if ((!chips || !out) && n) { set_error("null argument"); return UGSDR_ERR_ARG; }
    for (int64_t i = 0; i < n; ++i) {
        out[2 * i] = chips[i];
        out[2 * i + 1] = (int8_t)-chips[i];
    }
    return UGSDR_OK;
}

double ugsdr_adjust_sampling_rate(double fs, double target) {
    // keep the nearest integer-kHz divisor of fs; fall back to fs when the
    // divisor loop runs out without getting worse
    double best = fs;
    const size_t lim = (size_t)(2 * fs / target);
    for (size_t i = 1; i < lim; ++i) {
        const double cand = fs / (double)i;
        if (std::fmod(cand / 1e3, 1) != 0.0) continue;
        if (std::fabs(target - cand) > std::fabs(target - best)) return best;
        best = cand;
    }
    return fs;
}

// ------------------------------------------------------------------ replicas
static int replica_store(ugsdr_hip_ctx* ctx, const int8_t* chips, int64_t n_chips, int64_t repeats,
                         const float* samples, int64_t n_fft, int precision, uint64_t* key) {
    if (!ctx || !key) { set_error("null argument"); return UGSDR_ERR_ARG; }
    if (n_fft <= 0 || n_fft > (1 << 22)) { set_error("bad replica length"); return UGSDR_ERR_ARG; }
    if (precision != UGSDR_PREC_FP32 && precision != UGSDR_PREC_FP16 && precision != UGSDR_PREC_BF16) {
        set_error("unknown precision");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    Replica rep;
    rep.n_fft = n_fft;
    rep.precision = UGSDR_PREC_FP32;
    rep.repeats = chips ? repeats : 0;
    UGSDR_HIP_TRY(hipMalloc(&rep.spec, (size_t)n_fft * sizeof(float2)));
    int rc = replica_spectrum(ctx, chips, n_chips, repeats, samples, n_fft, static_cast<float2*>(rep.spec));
    if (rc) {
        (void)hipFree(rep.spec);
        return rc;
    }
    if (precision != UGSDR_PREC_FP32) {
        rc = replica_reduce(ctx, rep, precision);
        if (rc) {
            (void)hipFree(rep.spec);
            return rc;
        }
    } else if (chips && repeats > 1 && n_fft % repeats == 0) {
        // whole code periods: one period's spectrum for the periodic-fold
        // search (Upsampler::Transform of one period: the long replica's
        // first n_fft / repeats samples, the same chip indices)
        const int64_t L = n_fft / repeats;
        UGSDR_HIP_TRY(hipMalloc(&rep.spec_short, (size_t)L * sizeof(float2)));
        rc = replica_spectrum(ctx, chips, n_chips, 1, nullptr, L, static_cast<float2*>(rep.spec_short));
        if (rc) {
            (void)hipFree(rep.spec);
            (void)hipFree(rep.spec_short);
            return rc;
        }
        rep.period = L;
    }
    const bool fp32 = precision == UGSDR_PREC_FP32;
    if (chips && n_fft % repeats == 0 && (n_fft / repeats) % n_chips == 0 &&
        (fp32 ? upsampled_code_supported(n_chips, n_fft / repeats / n_chips)
              : repeats == 1 && reduced_ucode_supported(n_chips, n_fft / n_chips))) {
        // one period upsampled U times: the chips' own spectrum for the
        // upsampled-code search (16-bit for a reduced-precision replica)
        rc = hipMalloc(&rep.code_spec, (size_t)n_chips * sizeof(float2)) == hipSuccess
                 ? replica_spectrum(ctx, chips, n_chips, 1, nullptr, n_chips, static_cast<float2*>(rep.code_spec))
                 : (set_error("hipMalloc failed"), UGSDR_ERR_HIP);
        if (!rc && fp32) rc = code_spectrum_layout(ctx, n_chips, static_cast<float2*>(rep.code_spec));
        rep.code_len = n_chips;
        if (!rc && !fp32) rc = code_spec_reduce(ctx, rep, precision);
        if (rc) {
            (void)hipFree(rep.spec);
            if (rep.spec_short) (void)hipFree(rep.spec_short);
            if (rep.code_spec) (void)hipFree(rep.code_spec);
            return rc;
        }
        rep.code_len = n_chips;
    }
    *key = ctx->next_key++;
    ctx->replicas[*key] = rep;
    return UGSDR_OK;
}

int ugsdr_hip_replica_put(ugsdr_hip_ctx* ctx, const int8_t* chips, int64_t n_chips, int64_t repeats, int64_t n_fft,
                          int precision, uint64_t* key) {
    if (!chips || n_chips <= 0 || repeats <= 0) { set_error("bad chips"); return UGSDR_ERR_ARG; }
    return replica_store(ctx, chips, n_chips, repeats, nullptr, n_fft, precision, key);
}

int ugsdr_hip_replica_put_samples(ugsdr_hip_ctx* ctx, const float* samples, int64_t n_fft, int precision,
                                  uint64_t* key) {
    if (!samples) { set_error("null samples"); return UGSDR_ERR_ARG; }
    return replica_store(ctx, nullptr, 0, 1, samples, n_fft, precision, key);
}

int ugsdr_hip_replica_drop(ugsdr_hip_ctx* ctx, uint64_t key) {
    if (!ctx) { set_error("null context"); return UGSDR_ERR_ARG; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->replicas.find(key);
    if (it == ctx->replicas.end()) { set_error("unknown replica key"); return UGSDR_ERR_STATE; }
    if (it->second.users > 0) {
        set_error("replica in use by queued or running server requests");
        return UGSDR_ERR_STATE;
    }
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    UGSDR_HIP_TRY(hipFree(it->second.spec));
    if (it->second.spec_short) UGSDR_HIP_TRY(hipFree(it->second.spec_short));
    if (it->second.code_spec) UGSDR_HIP_TRY(hipFree(it->second.code_spec));
    ctx->replicas.erase(it);
    return UGSDR_OK;
}

int ugsdr_hip_replica_get(ugsdr_hip_ctx* ctx, uint64_t key, float* spectrum, int64_t n_fft) {
    if (!ctx || !spectrum) { set_error("null argument"); return UGSDR_ERR_ARG; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->replicas.find(key);
    if (it == ctx->replicas.end()) { set_error("unknown replica key"); return UGSDR_ERR_STATE; }
    if (it->second.n_fft != n_fft) { set_error("length mismatch"); return UGSDR_ERR_SIZE; }
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return replica_read(ctx, it->second, spectrum);
}

}  // extern "C"
