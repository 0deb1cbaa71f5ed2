// DFE channel step (Channel::GetSeveralEpochs, src/dfe/dfe.hpp:192-203):
// decode the recorded epochs, mix with the channel's running phase, resample,
// all in one pass of the front kernel over the recording's bytes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "ctx.hpp"
#include "formats.hpp"
#include "internal.hpp"

using namespace ugsdr_hip;

struct ugsdr_hip_dfe {
    ugsdr_hip_ctx* ctx = nullptr;
    int fmt = 0;
    double fs_in = 0, fs_out = 0;
    double f = 0;      // mixer frequency as given (the phase advance uses it unfolded)
    double phase = 0;  // running mixer phase, radians
    int spm_in = 0, epoch_bytes = 0;
    int64_t spm_out = 0, up = 1, down = 1;
    DevBuf in, out;    // staging for host buffers
};

static int64_t gcd_i64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static bool whole_khz(double fs) { return fs >= 1e3 && fs < 1e12 && std::fmod(fs, 1e3) == 0.0; }

extern "C" {

int ugsdr_hip_dfe_create(ugsdr_hip_ctx* ctx, int fmt, double fs_in, double fs_out, double mixer_frequency,
                         double phase, ugsdr_hip_dfe** out) {
    UGSDR_CHECK_ARG(ctx && out, "bad arguments");
    UGSDR_CHECK_ARG(whole_khz(fs_in) && whole_khz(fs_out), "DFE rates must be whole numbers of kHz");
    UGSDR_CHECK_ARG(std::isfinite(mixer_frequency) && std::isfinite(phase), "bad mixer parameters");
    int spm, eb;
    if (int rc = format_geometry(fmt, fs_in, spm, eb)) return rc;
    auto* ch = new ugsdr_hip_dfe;
    ch->ctx = ctx;
    ch->fmt = fmt;
    ch->fs_in = fs_in;
    ch->fs_out = fs_out;
    ch->f = mixer_frequency;
    ch->phase = phase;
    ch->spm_in = (int)(fs_in / 1e3);
    ch->epoch_bytes = eb;
    ch->spm_out = (int64_t)(fs_out / 1e3);
    const int64_t a = (int64_t)fs_in, b = (int64_t)fs_out;
    if (a != b) {
        const int64_t g = gcd_i64(a, b);
        ch->up = b / g;    // lcm / fs_in
        ch->down = a / g;  // lcm / fs_out
    }
    *out = ch;
    return UGSDR_OK;
}

int ugsdr_hip_dfe_process(ugsdr_hip_dfe* ch, const void* in, int in_on_device, int64_t n_epochs, float* out_c32,
                          int out_on_device) {
    UGSDR_CHECK_ARG(ch && in && out_c32 && n_epochs > 0, "bad arguments");
    const int64_t n_in = n_epochs * ch->spm_in, n_out = n_epochs * ch->spm_out;
    UGSDR_CHECK_ARG(n_in < ((int64_t)1 << 31) && n_out < ((int64_t)1 << 31), "at most 2^31 - 1 samples a call");
    const int64_t bytes = format_bytes(ch->fmt, n_in, ch->spm_in, ch->epoch_bytes);
    if (!in_on_device && ch->fmt == UGSDR_FMT_BBP_DDC)
        if (int rc = verify_bbp_headers(in, n_epochs, ch->epoch_bytes)) return rc;
    ugsdr_hip_ctx* ctx = ch->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const void* src = in;
    if (!in_on_device) {
        UGSDR_HIP_TRY(ch->in.reserve(bytes));
        UGSDR_HIP_TRY(hipMemcpyAsync(ch->in.ptr, in, bytes, hipMemcpyHostToDevice, st));
        src = ch->in.ptr;
    }
    float2* dst = reinterpret_cast<float2*>(out_c32);
    if (!out_on_device) {
        UGSDR_HIP_TRY(ch->out.reserve((size_t)n_out * sizeof(float2)));
        dst = ch->out.as<float2>();
    }
    // Mixer::Translate in-place overload (mixer.hpp:33-41): phase reduced to
    // [0, 2 pi), no-op at f == 0, f < 0 folded to fs + f
    double ph = std::fmod(ch->phase, 2 * M_PI);
    if (ph < 0) ph += 2 * M_PI;
    const double f = ch->f < 0 ? ch->fs_in + ch->f : ch->f;
    FrontArgs fa{};
    fa.src.p = src;
    fa.src.fmt = ch->fmt;
    fa.src.spm = ch->spm_in;
    fa.src.epoch_bytes = ch->epoch_bytes;
    fa.n_in = n_in;
    fa.n_out = n_out;
    fa.apply_tone = ch->f != 0.0 ? 1 : 0;
    fa.f_turns = f / ch->fs_in;
    fa.ph_turns = ph / (2 * M_PI);
    fa.up = ch->up;
    fa.down = ch->down;
    fa.out = dst;
    if (int rc = launch_front_args(fa, st)) return rc;
    if (!out_on_device)
        UGSDR_HIP_TRY(hipMemcpyAsync(out_c32, dst, (size_t)n_out * sizeof(float2), hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    // member Translate (mixer.hpp:53-60): the phase moves on by the fraction of
    // a cycle the tone turns in one millisecond, once per call
    double m = std::fmod(ch->f / 1000.0, 1.0);
    if (m < 0) m += 1;
    ch->phase += 2 * M_PI * m;
    return UGSDR_OK;
}

int ugsdr_hip_dfe_phase(ugsdr_hip_dfe* ch, double* phase) {
    UGSDR_CHECK_ARG(ch && phase, "bad arguments");
    *phase = ch->phase;
    return UGSDR_OK;
}

int ugsdr_hip_dfe_destroy(ugsdr_hip_dfe* ch) {
    if (!ch) return UGSDR_OK;
    {
        std::lock_guard<std::mutex> lk(ch->ctx->mu);
        (void)hipSetDevice(ch->ctx->device);
        ch->in.release();
        ch->out.release();
    }
    delete ch;
    return UGSDR_OK;
}

}  // extern "C"
