// RAII owner of a ugsdr_hip_ctx and the error bridge of the C++ surface:
// every failing C-ABI call is rethrown as std::runtime_error, the reference's
// error convention (e.g. src/math/reshape_and_sum.hpp:16-17).
#pragma once

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

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

}  // namespace ugsdr_hip
