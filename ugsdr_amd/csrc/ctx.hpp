// Context, error plumbing and device buffer management shared by the
// library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ugsdr_hip.h"

namespace ugsdr_hip {

void set_error(const std::string& msg);

#define UGSDR_HIP_TRY(expr)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ::ugsdr_hip::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            return UGSDR_ERR_HIP;                                                           \
        }                                                                                   \
    } while (0)

#define UGSDR_CHECK_ARG(cond, msg)                  \
    do {                                            \
        if (!(cond)) {                              \
            ::ugsdr_hip::set_error(msg);            \
            return UGSDR_ERR_ARG;                   \
        }                                           \
    } while (0)

// Raises a kernel's dynamic LDS limit to at least `bytes` on the current
// device; the attribute is set once per (kernel, device) and remembered.
hipError_t set_lds_limit(const void* kernel, size_t bytes);

// grow-only device buffer
// Set by a thread while it captures a stream into a graph: a DevBuf that
// would have to grow then fails instead (no free or allocation inside a
// capture), and the caller runs the steps without a graph.
inline thread_local bool g_capturing = false;

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (g_capturing) return hipErrorStreamCaptureUnsupported;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            bytes = 0;
        }
        size_t want = n < 4096 ? 4096 : n;
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// page-locked host memory (hipHostMalloc), grown on demand
struct PinnedBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipHostMalloc(&ptr, n, hipHostMallocDefault);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    PinnedBuf(PinnedBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    ~PinnedBuf() {
        if (ptr) (void)hipHostFree(ptr);
    }
};

struct Replica {
    void* spec = nullptr;  // n_fft complex (float2, or half2/bf16x2 for reduced)
    int64_t n_fft = 0;
    int precision = UGSDR_PREC_FP32;
    float scale = 1.0f;    // reduced precision: stored = spectrum * scale
    int64_t users = 0;     // queued or running server requests that read it
    // a replica of `repeats` whole code periods (replica_put with n_fft a
    // multiple of repeats, FP32): the period in samples and conj(DFT) of one
    // period, for the periodic-fold search (otherwise 0 / null)
    int64_t period = 0;
    void* spec_short = nullptr;
    // the replica is one code period of code_len chips upsampled by an
    // integer factor (repeats == 1, or one period of a periodic replica):
    // conj(DFT) of the code_len chips themselves, for the upsampled-code
    // search (otherwise null)
    int64_t repeats = 1;
    void* code_spec = nullptr;   // reduced precision: 16-bit pairs, times code_scale
    int64_t code_len = 0;
    float code_scale = 1.0f;
};

// device buffers one acquisition request holds while in flight
struct AcqSlot {
    DevBuf iq, sig, classes, tables, rows, results, surf, fwd_scratch;
    // the recorded samples the class transforms read directly when the front
    // end only decodes them (int16 or float I/Q at the acquisition rate, no
    // tone): set by launch_front, null when it filled sig
    const void* front_src = nullptr;
    int front_fmt = 0;
    // periodic-fold searches: the class spectra of every code-period sub-block
    DevBuf classes_sub;
    // upsampled-code searches: the PRN-independent spectra Z per (Doppler bin, block)
    DevBuf zbuf;
    // reduced precision: 16-bit class spectra and their per-block scales
    DevBuf classes_r, block_scale;
    std::vector<char> host_tables;   // kept alive until the copies complete
    void release_all() {
        for (DevBuf* d : {&iq, &sig, &classes, &tables, &rows, &results, &surf, &fwd_scratch, &classes_r, &block_scale,
                          &classes_sub, &zbuf})
            d->release();
    }
    int64_t bytes() const {
        int64_t b = 0;
        for (const DevBuf* d : {&iq, &sig, &classes, &tables, &rows, &results, &surf, &fwd_scratch, &classes_r,
                                &block_scale, &classes_sub, &zbuf})
            b += (int64_t)d->bytes;
        return b;
    }
};

}  // namespace ugsdr_hip

struct ugsdr_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::map<uint64_t, ugsdr_hip::Replica> replicas;
    uint64_t next_key = 1;
    // scratch
    ugsdr_hip::DevBuf iq, sig, classes, scratch, rows, results, surf, tables, aux, aux2;
    ugsdr_hip::AcqSlot slot;  // buffers of the synchronous acquire path
    // the synchronous acquire path's results (and peaks) land here first: a
    // copy to page-locked memory returns sooner than one to pageable memory
    ugsdr_hip::PinnedBuf results_host;
    // hipGraph of the last repeated request shape (ugsdr_hip_acquire): the
    // shape is captured on its second consecutive call and replayed after
    struct GraphCache {
        bool enabled = true;
        std::string key, pending;
        hipGraphExec_t exec = nullptr;
        int64_t replays = 0;
        void drop() {
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
            key.clear();
        }
    } graph;
    // tracking replica tables (ugsdr_hip_track), kept while the channels'
    // codes and the sample rate stay the same between calls
    struct TrackTables {
        ugsdr_hip::DevBuf buf;
        std::vector<int8_t> codes;
        int64_t spm = 0, period = 0, len3p = 0;
        int n_ch = 0;
    } track;
    // upsampled-code searches: conj of the U-sample boxcar's N-point
    // spectrum, per (N, U) (key N * 64 + U)
    std::map<int64_t, ugsdr_hip::DevBuf> boxcar;
    // and max over k' < N / U of sum_u |H[k' + u N / U]| (bounds |Z| by the
    // class blocks' max |X|; the reduced form scales Z with it)
    std::map<int64_t, float> boxcar_bound;
    // the form of the last ugsdr_hip_acquire (ugsdr_hip_last_search_form)
    int64_t last_n = 0;
    int last_periods = 0, last_ucode = 0;
    // autotuner: transform length -> correlation workgroup size
    std::map<int64_t, int> tuned;
    bool autotune = true;
    // stage timing (ugsdr_hip_ctx_set_profiling)
    bool profile = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float last_ms[4] = {0, 0, 0, 0};
    void mark(int i) {
        if (profile && ev[i]) (void)hipEventRecord(ev[i], stream);
    }
    int64_t device_bytes() const;
};
