// ugsdr_hip_acquire_noalloc: the fused FP32 acquisition of ugsdr_hip_acquire
// for callers that cannot allocate device memory.  Every device buffer of
// the search (tables, uploaded IQ, the front end's output, the class
// spectra, the four-step transforms' workspace, the row statistics, the
// correlation kernels' partial surfaces and the results) is carved out of
// one file-scope device array of fixed size, so a call runs no hipMalloc,
// hipFree, graph instantiation or autotuning.  The kernels are the ones
// ugsdr_hip_acquire launches; they get the array's device address as their
// buffer pointers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "acq.hpp"
#include "ctx.hpp"
#include "internal.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // the stated maximum: 64 MiB per device
constexpr int kMaxDevices = 64;

// The device-side workspace.  Its name is a host-side stand-in in host code:
// only the address hipGetSymbolAddress returns for a device may be handed to
// kernels or copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_init_mu;
void* g_scratch_ptr[kMaxDevices] = {};
// one call at a time per device: every call on a device shares its array
std::mutex g_device_mu[kMaxDevices];

// bump allocation inside the array, 256-byte aligned
struct Carver {
    char* base;
    size_t cap;
    size_t used = 0;
    bool take(ugsdr_hip::DevBuf& b, size_t n) {
        const size_t off = (used + 255) & ~size_t(255);
        if (off > cap || n > cap - off) return false;
        b.ptr = base + off;
        b.bytes = n;
        used = off + n;
        return true;
    }
    size_t left() const {
        const size_t off = (used + 255) & ~size_t(255);
        return off >= cap ? 0 : cap - off;
    }
};

}  // namespace

namespace ugsdr_hip {

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Resolves the array's address on `device` (current) once per device; called
// by ugsdr_hip_ctx_create.
int noalloc_scratch_init(int device) {
    UGSDR_CHECK_ARG(device >= 0 && device < kMaxDevices, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_scratch_ptr[device]) return UGSDR_OK;
    void* p = nullptr;
    UGSDR_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("hipGetSymbolAddress returned no address for the no-alloc workspace");
        return UGSDR_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
    return UGSDR_OK;
}

static int acquire_noalloc(ugsdr_hip_ctx* ctx, char* base, const ugsdr_acq_request* r, ugsdr_acq_result* out,
                           float* surfaces, float* peaks) {
    UGSDR_CHECK_ARG(r && out, "null request or result");
    if (r->precision != UGSDR_PREC_FP32) {
        set_error("the no-alloc entry point runs the FP32 plain form only");
        return UGSDR_ERR_UNSUPPORTED;
    }
    const int H = r->hypotheses;
    UGSDR_CHECK_ARG(H <= 1 || (r->n_prn > 0 && r->n_prn % H == 0), "n_prn must be a multiple of the hypothesis count");
    AcqPlan pl;
    // the upsampled-code search's boxcar table is allocated on first use:
    // this entry point runs the other search forms only
    if (int rc = make_acq_plan(ctx, r, pl, false)) return rc;
    const bool keep_surf = surfaces || peaks;
    const int cus = device_cu_count(ctx->device);
    // the buffers ugsdr_hip_acquire's slot would hold, in the sizes its
    // enqueue steps ask for (a DevBuf already that large is used as is)
    AcqSlot s;
    DevBuf corr_scratch;
    hipStream_t st = ctx->stream;
    // whatever path leaves this function, the stream has drained first: the
    // host tables of s are the source of an asynchronous upload, and the next
    // call on this device reuses the workspace
    struct Drain {
        hipStream_t st;
        ~Drain() { (void)hipStreamSynchronize(st); }
    } drain{st};
    Carver cv{base, kScratchBytes};
    bool ok = cv.take(s.tables, pl.table_bytes) && cv.take(s.sig, (size_t)pl.need * sizeof(float2)) &&
              cv.take(s.rows, (size_t)pl.P * pl.D * sizeof(RowStat)) &&
              cv.take(s.classes, (size_t)pl.K * pl.F * pl.N * sizeof(float2)) &&
              cv.take(s.results, (size_t)pl.P * sizeof(ugsdr_acq_result) + (peaks ? (size_t)pl.P * pl.L * 4 : 0));
    if (ok && !r->iq_on_device) ok = cv.take(s.iq, (size_t)pl.in_bytes);
    if (ok && keep_surf) ok = cv.take(s.surf, (size_t)pl.P * pl.D * pl.L * sizeof(float));
    if (ok && pl.Tsub > 1) ok = cv.take(s.classes_sub, (size_t)pl.K * pl.Tsub * pl.F * pl.N * sizeof(float2));
    if (ok && pl.fwd_scratch_needed()) ok = cv.take(s.fwd_scratch, (size_t)pl.K * pl.Tsub * pl.F * pl.N * sizeof(float2));
    // correlation workspace: the three-stage kernel's partial surfaces, or one
    // four-step slice per workgroup (fewer workgroups when the array is short:
    // every row is still computed whole by one workgroup)
    int corr_cus = cus;
    if (ok) {
        if (pl.small) {
            const size_t need = corr_scratch_bytes(pl, cus);
            ok = need == 0 || cv.take(corr_scratch, need);
        } else {
            const size_t per_wg = (2 * (size_t)pl.N + (size_t)(pl.L + 1) / 2) * sizeof(float2);
            const size_t fit = cv.left() / per_wg;
            corr_cus = (int)std::min<size_t>((size_t)cus, fit);
            ok = corr_cus >= 1 && cv.take(corr_scratch, per_wg * (size_t)corr_cus);
        }
    }
    if (!ok) {
        set_error("the search needs more than the no-alloc workspace of " + std::to_string(kScratchBytes >> 20) +
                  " MiB");
        return UGSDR_ERR_NOMEM;
    }
    // the tuned workgroup size when this context has one, else the default:
    // the no-alloc path never autotunes (the autotuner allocates)
    const auto it = ctx->tuned.find(pl.N);
    const int threads = it != ctx->tuned.end() ? it->second : default_threads(pl.N);
    if (int rc = enqueue_front(pl, r, s, st, keep_surf, FrontStep::All)) return rc;
    if (int rc = enqueue_classes(pl, s, st)) return rc;
    if (int rc = enqueue_corr({&pl}, {&s}, st, corr_scratch, pl.small ? cus : corr_cus, threads)) return rc;
    if (int rc = enqueue_final(pl, s, st, peaks != nullptr)) return rc;
    std::vector<ugsdr_acq_result> all;
    ugsdr_acq_result* dst = out;
    if (H > 1) {
        all.resize((size_t)pl.P);
        dst = all.data();
    }
    UGSDR_HIP_TRY(hipMemcpyAsync(dst, s.results.ptr, (size_t)pl.P * sizeof(ugsdr_acq_result), hipMemcpyDeviceToHost,
                                 st));
    if (peaks)
        UGSDR_HIP_TRY(hipMemcpyAsync(peaks, s.results.as<char>() + (size_t)pl.P * sizeof(ugsdr_acq_result),
                                     (size_t)pl.P * pl.L * 4, hipMemcpyDeviceToHost, st));
    if (surfaces)
        UGSDR_HIP_TRY(hipMemcpyAsync(surfaces, s.surf.ptr, (size_t)pl.P * pl.D * pl.L * 4, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    if (H > 1) reduce_hypotheses(all.data(), pl.P, H, r->threshold, out);
    return UGSDR_OK;
}

}  // namespace ugsdr_hip

extern "C" int ugsdr_hip_acquire_noalloc(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* req, ugsdr_acq_result* out,
                                         float* surfaces, float* peaks) {
    if (!ctx) {
        ugsdr_hip::set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    void* base = nullptr;
    if (ctx->device >= 0 && ctx->device < kMaxDevices) {
        std::lock_guard<std::mutex> init(g_init_mu);
        base = g_scratch_ptr[ctx->device];
    }
    if (!base) {
        ugsdr_hip::set_error("no-alloc workspace not resolved for this device");
        return UGSDR_ERR_STATE;
    }
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> dev(g_device_mu[ctx->device]);
    return ugsdr_hip::acquire_noalloc(ctx, static_cast<char*>(base), req, out, surfaces, peaks);
}
