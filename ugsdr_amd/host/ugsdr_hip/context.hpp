// RAII owner of a ugsdr_hip_ctx and the error bridge of the C++ surface:
// every failing C-ABI call is rethrown as std::runtime_error, the reference's
// error convention (e.g. src/math/reshape_and_sum.hpp:16-17).
#pragma once

#include <complex>
#include <cstddef>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ugsdr_hip.h"

namespace ugsdr_hip {

inline void check(int rc) {
    if (rc != UGSDR_OK) throw std::runtime_error(ugsdr_hip_last_error());
}

class HipContext {
public:
    explicit HipContext(int device = 0) { check(ugsdr_hip_ctx_create(device, &ctx_)); }
    ~HipContext() { ugsdr_hip_ctx_destroy(ctx_); }
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;

    ugsdr_hip_ctx* get() const { return ctx_; }

    // process-wide context on device 0 used by the static CRTP operators,
    // mirroring the reference's implicit per-thread IPP state
    static HipContext& Default() {
        static HipContext* ctx = new HipContext(0);
        return *ctx;
    }

private:
    ugsdr_hip_ctx* ctx_ = nullptr;
};

// SignalParametersBase::GetPartialSignal (src/signal_parameters.hpp:140-237)
// from the start of a recording held in memory: n samples of fmt
// (UGSDR_FMT_*, packed BbpDdc / Nt1065 included) decoded on the GPU.  Throws
// std::runtime_error on a bad format or rate and, for BbpDdc, on an epoch ms
// counter gap ("Header ms count mismatch", signal_parameters.hpp:41-42).
inline std::vector<std::complex<float>> GetPartialSignal(const void* bytes, int fmt, std::size_t n, double fs,
                                                         HipContext& ctx = HipContext::Default()) {
    std::vector<std::complex<float>> out(n);
    check(ugsdr_hip_unpack(ctx.get(), bytes, fmt, (int64_t)n, fs, reinterpret_cast<float*>(out.data())));
    return out;
}

}  // namespace ugsdr_hip
