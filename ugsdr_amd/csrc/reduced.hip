// Reduced-precision storage of replica spectra (FP16 / BF16) and the
// reduced-precision acquisition entry.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "internal.hpp"

namespace ugsdr_hip {

__global__ void to_half2_kernel(const float2* in, __half2* out, int64_t n, float scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __floats2half2_rn(in[i].x * scale, in[i].y * scale);
}
__global__ void to_bf162_kernel(const float2* in, __hip_bfloat162* out, int64_t n, float scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __hip_bfloat162(__float2bfloat16(in[i].x * scale), __float2bfloat16(in[i].y * scale));
}
__global__ void from_half2_kernel(const __half2* in, float2* out, int64_t n, float inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = __half22float2(in[i]);
        out[i] = make_float2(v.x * inv, v.y * inv);
    }
}
__global__ void from_bf162_kernel(const __hip_bfloat162* in, float2* out, int64_t n, float inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = __bfloat1622float2(in[i]);
        out[i] = make_float2(v.x * inv, v.y * inv);
    }
}

int replica_reduce(ugsdr_hip_ctx* ctx, Replica& rep, int precision) {
    // Parseval: sum |C_k|^2 = N * sum c^2 = N^2 for +-1 replicas, so the RMS
    // bin magnitude is sqrt(N); scaling by 1/sqrt(N) centres it at 1.
    const float scale = (float)(1.0 / std::sqrt((double)rep.n_fft));
    void* dst = nullptr;
    UGSDR_HIP_TRY(hipMalloc(&dst, (size_t)rep.n_fft * 4));
    const int blocks = (int)std::min<int64_t>((rep.n_fft + 255) / 256, 4096);
    if (precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(to_half2_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.spec), static_cast<__half2*>(dst), rep.n_fft, scale);
    else
        hipLaunchKernelGGL(to_bf162_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const float2*>(rep.spec), static_cast<__hip_bfloat162*>(dst), rep.n_fft,
                           scale);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    UGSDR_HIP_TRY(hipFree(rep.spec));
    rep.spec = dst;
    rep.precision = precision;
    rep.scale = scale;
    return UGSDR_OK;
}

int replica_read(ugsdr_hip_ctx* ctx, const Replica& rep, float* out) {
    if (rep.precision == UGSDR_PREC_FP32) {
        UGSDR_HIP_TRY(hipMemcpyAsync(out, rep.spec, (size_t)rep.n_fft * 8, hipMemcpyDeviceToHost, ctx->stream));
        UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return UGSDR_OK;
    }
    UGSDR_HIP_TRY(ctx->aux.reserve((size_t)rep.n_fft * 8));
    const int blocks = (int)std::min<int64_t>((rep.n_fft + 255) / 256, 4096);
    if (rep.precision == UGSDR_PREC_FP16)
        hipLaunchKernelGGL(from_half2_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const __half2*>(rep.spec), ctx->aux.as<float2>(), rep.n_fft, 1.0f / rep.scale);
    else
        hipLaunchKernelGGL(from_bf162_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<const __hip_bfloat162*>(rep.spec), ctx->aux.as<float2>(), rep.n_fft,
                           1.0f / rep.scale);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, ctx->aux.ptr, (size_t)rep.n_fft * 8, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

int acquire_reduced(ugsdr_hip_ctx*, const ugsdr_acq_request*, ugsdr_acq_result*, float*, float*) {
    set_error("reduced-precision acquisition is not built yet");
    return UGSDR_ERR_UNSUPPORTED;
}

}  // namespace ugsdr_hip
