// Anti-jamming front end: the short-time Fourier transform of
// src/math/stft.hpp / ipp_stft.hpp, the detector of
// src/antijamming/jamming_detection.hpp and the suppression of
// src/antijamming/jse.hpp, with the per-column work on the GPU.
//
//   stft_kernel     one 16-lane group per column: 128 windowed samples, a
//                   128-point DFT as 16 DFT-8s, a twiddle, 8 DFT-16s through
//                   LDS; with DETECT, the eight strongest bins of the column
//                   (each nulled +-3 after it is taken) and the mean and
//                   n-1 sigma of the nulled magnitudes, in FP64
//   istft_kernel    the inverse of each column after zeroing the cells at or
//                   above the absolute threshold, in place
//   ola_kernel      overlap-add of the inverse columns at stride 32, in
//                   column order, cut or zero-padded to the output length
//   notch_kernel    zeroes a run of DFT bins modulo n (the narrowband notch)
//
// The per-column candidates go to the host, which runs AnalyzeStftData (a
// few hundred bytes of decisions) and picks the suppression.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

namespace {

constexpr int kWin = 128, kOverlap = 96, kStride = kWin - kOverlap;
constexpr int kCand = 8;         // max_inreference_cnt (jamming_detection.hpp:70)
constexpr int kDetNull = 3;      // the detector's null_window_size (:71)
constexpr int kColsPerBlock = 16;
constexpr int kLdsPitch = 17;    // odd pitch: lanes of a group hit different banks

inline int grid_of(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 1 << 20);
}

// X[k1 + 8 k2] = sum_t W16^(t k2) W128^(t k1) sum_m x[t + 16 m] W8^(m k1)
// for the group's column: lane t (0..15) runs the DFT-8 over m, lanes
// k1 = 0..7 the DFT-16 over t.  On return lanes 0..7 hold v[k2] = X[k1 + 8 k2].
template <bool INV>
__device__ __forceinline__ void dft128(float2 (&v)[16], float2* lds, int t) {
    Dft<8, INV>::run(v);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
        // W128^(+-t k1) = e^(-+j pi t k1 / 64)
        float s, c;
        sincospif((INV ? 1.f : -1.f) * (float)(t * k1) / 64.f, &s, &c);
        lds[k1 * kLdsPitch + t] = k1 == 0 ? v[0] : cmul(v[k1], make_float2(c, s));
    }
    __syncthreads();
    const int k1 = t & 7;
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = lds[k1 * kLdsPitch + j];
    Dft<16, INV>::run(v);
}

template <int WINDOWED_DETECT>
__global__ __launch_bounds__(256) void stft_kernel(const float2* x, int64_t ncol, const float* window,
                                                   float2* cols, int* cand_bin, float* cand_val, double* sigma) {
    __shared__ float2 lds[kColsPerBlock][8 * kLdsPitch];
    const int g = threadIdx.x >> 4, t = threadIdx.x & 15;
    const int64_t c = (int64_t)blockIdx.x * kColsPerBlock + g;
    const bool live = c < ncol;
    const int64_t cc = live ? c : ncol - 1;   // a spare group repeats the last column and writes nothing
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = t + 16 * m;
        const float2 s = x[cc * kStride + i];
        const float w = window[i];
        v[m] = make_float2(s.x * w, s.y * w);
    }
    dft128<false>(v, lds[g], t);
    const int k1 = t & 7;
    const bool owner = live && t < 8;
    if (owner) {
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) cols[cc * kWin + k1 + 8 * k2] = v[k2];
    }
    if (!WINDOWED_DETECT) return;
    // IntermediateDetection (jamming_detection.hpp:115-137) on lanes 0..7 of
    // the group (lanes 8..15 mirror them and write nothing)
    float mag[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) mag[k2] = sqrtf(v[k2].x * v[k2].x + v[k2].y * v[k2].y);
    for (int r = 0; r < kCand; ++r) {
        float best = mag[0];
        int bi = k1;
#pragma unroll
        for (int k2 = 1; k2 < 16; ++k2)
            if (mag[k2] > best) {
                best = mag[k2];
                bi = k1 + 8 * k2;
            }
#pragma unroll
        for (int off = 4; off > 0; off >>= 1) {
            const float ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (owner && k1 == 0) {
            cand_bin[cc * kCand + r] = bi;
            cand_val[cc * kCand + r] = best;
        }
        // NullRegion(index, 3, magnitude): the seven bins around it, circularly
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int d = (k1 + 8 * k2 - bi + kWin) & (kWin - 1);
            if (d <= kDetNull || d >= kWin - kDetNull) mag[k2] = 0.f;
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) s += (double)mag[k2];
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const double mean = s / kWin;
    double q = 0.0;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) q += ((double)mag[k2] - mean) * ((double)mag[k2] - mean);
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if (owner && k1 == 0) sigma[cc] = sqrt(q / (kWin - 1));
}

// inverse DFT (1/128) of each column in place; cells with |X| >= threshold
// are zeroed first (FilterWideband, jse.hpp:71-87)
__global__ __launch_bounds__(256) void istft_kernel(float2* cols, int64_t ncol, float threshold) {
    __shared__ float2 lds[kColsPerBlock][8 * kLdsPitch];
    const int g = threadIdx.x >> 4, t = threadIdx.x & 15;
    const int64_t c = (int64_t)blockIdx.x * kColsPerBlock + g;
    const bool live = c < ncol;
    const int64_t cc = live ? c : ncol - 1;
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        float2 s = cols[cc * kWin + t + 16 * m];
        if (!(hypotf(s.x, s.y) < threshold)) s = make_float2(0.f, 0.f);
        v[m] = s;
    }
    dft128<true>(v, lds[g], t);
    const int k1 = t & 7;
    if (live && t < 8) {
        constexpr float inv = 1.f / kWin;
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) cols[cc * kWin + k1 + 8 * k2] = make_float2(v[k2].x * inv, v[k2].y * inv);
    }
}

// OverlapAndAdd (stft.hpp:60-72) into out[0, n): the columns covering sample
// m added in column order, zero past 128 + 32 (ncol - 1)
__global__ void ola_kernel(const float2* cols, int64_t ncol, float2* out, int64_t n) {
    const int64_t total = kWin + (ncol - 1) * kStride;
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        if (m < total) {
            const int64_t hi = std::min<int64_t>(m / kStride, ncol - 1);
            const int64_t lo = m >= kWin ? (m - kWin) / kStride + 1 : 0;
            for (int64_t i = lo; i <= hi; ++i) acc = cadd(acc, cols[i * kWin + (m - i * kStride)]);
        }
        out[m] = acc;
    }
}

__global__ void notch_kernel(float2* d, int64_t n, int64_t start, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        d[(start + i) % n] = make_float2(0.f, 0.f);
}

// IppStft's 4-term Blackman-Harris over N - 1, accumulated into a float term
// by term (ipp_stft.hpp:26-35), or SequentialStft's periodic Hann
// (stft.hpp:31-37)
void make_window(int kind, float* w) {
    for (int i = 0; i < kWin; ++i) {
        if (kind == UGSDR_STFT_BLACKMAN_HARRIS) {
            static const double a[4] = {0.3635819, 0.4891775, 0.1365995, 0.0106411};
            float acc = 0.f;  // float += double: the sum in double, rounded each term
            for (int j = 0; j < 4; ++j)
                acc = (float)((double)acc + a[j] * std::pow(-1.0, j) * std::cos(2.0 * j * M_PI * i / (kWin - 1)));
            w[i] = acc;
        } else {
            w[i] = (float)(0.5 * (1 - std::cos(2.0 * M_PI * (double)i / kWin)));
        }
    }
}

struct Internal {
    int bin;
    double frequency, jn;
};

// AnalyzeStftData (jamming_detection.hpp:139-190), its integer mean count,
// frequency-ordered map and size_t(size * 0.9) wideband rule included
std::vector<ugsdr_jammer> analyze(const std::vector<std::vector<Internal>>& inter) {
    constexpr double kProb = 0.9;
    std::vector<ugsdr_jammer> dst;
    size_t mean = 0;
    for (auto& el : inter) mean += el.size();
    mean /= inter.size();
    for (size_t i = 0; i < mean; ++i) {
        std::map<double, std::pair<double, Internal>> occ;
        double quantity = 0;
        for (auto& el : inter) {
            if (el.size() <= i) continue;
            auto& e = occ[el[i].frequency];
            e.first += 1;
            e.second.frequency = el[i].frequency;
            e.second.bin = el[i].bin;
            e.second.jn += el[i].jn;
            ++quantity;
        }
        for (auto& kv : occ) {
            kv.second.first /= quantity;
            kv.second.second.jn /= quantity;
        }
        size_t cnt_wideband = 0;
        const Internal* wb = &occ.begin()->second.second;
        for (auto& kv : occ) {
            const double p = kv.second.first;
            if (p >= kProb) {
                dst.push_back({UGSDR_JAMMER_NARROWBAND, kv.second.second.bin, kv.second.second.frequency,
                               kv.second.second.jn});
                break;
            }
            if (p >= kProb / (double)occ.size()) {
                ++cnt_wideband;
                wb = &kv.second.second;
            }
        }
        if (cnt_wideband >= (size_t)(occ.size() * kProb))
            dst.push_back({UGSDR_JAMMER_WIDEBAND, wb->bin, wb->frequency, wb->jn});
    }
    return dst;
}

int upload(ugsdr_hip_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
    UGSDR_HIP_TRY(b.reserve(bytes));
    UGSDR_HIP_TRY(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return UGSDR_OK;
}

int64_t columns_of(int64_t n) { return (n - kOverlap) / kStride; }

int stft_impl(ugsdr_hip_ctx* ctx, const float* in, int64_t n, int window, float* out) {
    UGSDR_CHECK_ARG(in && out && n >= kWin, "the STFT needs at least 128 samples");
    UGSDR_CHECK_ARG(window == UGSDR_STFT_BLACKMAN_HARRIS || window == UGSDR_STFT_HANN, "unknown window");
    const int64_t ncol = columns_of(n);
    float w[kWin];
    make_window(window, w);
    if (int rc = upload(ctx, ctx->tables, w, sizeof w)) return rc;
    if (int rc = upload(ctx, ctx->aux, in, (size_t)n * 8)) return rc;
    UGSDR_HIP_TRY(ctx->sig.reserve((size_t)ncol * kWin * 8));
    hipLaunchKernelGGL(stft_kernel<0>, dim3((unsigned)((ncol + kColsPerBlock - 1) / kColsPerBlock)), dim3(256), 0,
                       ctx->stream, ctx->aux.as<float2>(), ncol, ctx->tables.as<float>(), ctx->sig.as<float2>(),
                       nullptr, nullptr, nullptr);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, ctx->sig.ptr, (size_t)ncol * kWin * 8, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

int istft_impl(ugsdr_hip_ctx* ctx, const float* cols, int64_t ncol, float* out) {
    UGSDR_CHECK_ARG(cols && out && ncol > 0, "bad columns");
    if (int rc = upload(ctx, ctx->sig, cols, (size_t)ncol * kWin * 8)) return rc;
    const int64_t n = kWin + (ncol - 1) * kStride;
    UGSDR_HIP_TRY(ctx->aux.reserve((size_t)n * 8));
    hipLaunchKernelGGL(istft_kernel, dim3((unsigned)((ncol + kColsPerBlock - 1) / kColsPerBlock)), dim3(256), 0,
                       ctx->stream, ctx->sig.as<float2>(), ncol, INFINITY);
    hipLaunchKernelGGL(ola_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, ctx->sig.as<float2>(), ncol,
                       ctx->aux.as<float2>(), n);
    UGSDR_HIP_TRY(hipGetLastError());
    UGSDR_HIP_TRY(hipMemcpyAsync(out, ctx->aux.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

int jse_impl(ugsdr_hip_ctx* ctx, float* io, int64_t n, double fs, int window, int on_device, ugsdr_jammer* out,
             int32_t max_out, int32_t* n_out) {
    UGSDR_CHECK_ARG(io && n >= kWin && fs > 0, "the suppression needs at least 128 samples");
    UGSDR_CHECK_ARG(window == UGSDR_STFT_BLACKMAN_HARRIS || window == UGSDR_STFT_HANN, "unknown window");
    UGSDR_CHECK_ARG(max_out >= 0 && (max_out == 0 || out), "bad detection buffer");
    hipStream_t st = ctx->stream;
    const int64_t ncol = columns_of(n);
    float w[kWin];
    make_window(window, w);
    if (int rc = upload(ctx, ctx->tables, w, sizeof w)) return rc;
    float2* x = reinterpret_cast<float2*>(io);
    if (!on_device) {
        if (int rc = upload(ctx, ctx->aux, io, (size_t)n * 8)) return rc;
        x = ctx->aux.as<float2>();
    }
    UGSDR_HIP_TRY(ctx->sig.reserve((size_t)ncol * kWin * 8));
    const size_t b_bin = (size_t)ncol * kCand * 4, o_val = (b_bin + 255) / 256 * 256;
    const size_t o_sig = (o_val + b_bin + 255) / 256 * 256;
    UGSDR_HIP_TRY(ctx->results.reserve(o_sig + (size_t)ncol * 8));
    int* d_bin = ctx->results.as<int>();
    float* d_val = reinterpret_cast<float*>(ctx->results.as<char>() + o_val);
    double* d_sig = reinterpret_cast<double*>(ctx->results.as<char>() + o_sig);
    const unsigned blocks = (unsigned)((ncol + kColsPerBlock - 1) / kColsPerBlock);
    ctx->mark(0);
    hipLaunchKernelGGL(stft_kernel<1>, dim3(blocks), dim3(256), 0, st, x, ncol, ctx->tables.as<float>(),
                       ctx->sig.as<float2>(), d_bin, d_val, d_sig);
    UGSDR_HIP_TRY(hipGetLastError());
    ctx->mark(1);
    std::vector<int> bins((size_t)ncol * kCand);
    std::vector<float> vals((size_t)ncol * kCand);
    std::vector<double> sig((size_t)ncol);
    UGSDR_HIP_TRY(hipMemcpyAsync(bins.data(), d_bin, b_bin, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipMemcpyAsync(vals.data(), d_val, b_bin, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipMemcpyAsync(sig.data(), d_sig, (size_t)ncol * 8, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    // the candidates' J/N, those under 6 dropped in order (:130-136)
    std::vector<std::vector<Internal>> inter((size_t)ncol);
    for (int64_t c = 0; c < ncol; ++c)
        for (int r = 0; r < kCand; ++r) {
            const int b = bins[(size_t)c * kCand + r];
            const double jn = (double)vals[(size_t)c * kCand + r] / sig[(size_t)c];
            if (!(jn < 6.0)) inter[(size_t)c].push_back({b, b * fs / kWin, jn});
        }
    const std::vector<ugsdr_jammer> det = analyze(inter);
    if (n_out) *n_out = (int32_t)det.size();
    for (size_t i = 0; i < det.size() && (int32_t)i < max_out; ++i) out[i] = det[i];
    ctx->mark(2);
    if (!det.empty()) {
        const bool wideband = std::any_of(det.begin(), det.end(),
                                          [](const ugsdr_jammer& d) { return d.type == UGSDR_JAMMER_WIDEBAND; });
        if (!wideband) {
            // FilterInFreqencyDomain (jse.hpp:61-69): +-128 bins around each
            UGSDR_HIP_TRY(ctx->aux2.reserve((size_t)n * 8));
            if (int rc = dft_device(ctx, x, ctx->aux2.as<float2>(), n, 0)) return rc;
            for (const auto& d : det) {
                const int64_t idx = (int64_t)(size_t)(d.frequency / fs * (double)n);
                int64_t start = idx - 128 + n, finish = idx + 128;
                if (finish < start) finish += n;
                if (finish >= start)
                    hipLaunchKernelGGL(notch_kernel, dim3(grid_of(finish - start + 1)), dim3(256), 0, st, x, n,
                                       start, finish - start + 1);
            }
            if (int rc = dft_device(ctx, x, ctx->aux2.as<float2>(), n, 1)) return rc;
        } else {
            hipLaunchKernelGGL(istft_kernel, dim3(blocks), dim3(256), 0, st, ctx->sig.as<float2>(), ncol, 20.f);
            hipLaunchKernelGGL(ola_kernel, dim3(grid_of(n)), dim3(256), 0, st, ctx->sig.as<float2>(), ncol, x, n);
        }
        UGSDR_HIP_TRY(hipGetLastError());
    }
    ctx->mark(3);
    if (!on_device) UGSDR_HIP_TRY(hipMemcpyAsync(io, x, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    if (ctx->profile) {
        // STFT + detection, host analysis, suppression
        UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev[0], ctx->ev[1]));
        UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[1], ctx->ev[1], ctx->ev[2]));
        UGSDR_HIP_TRY(hipEventElapsedTime(&ctx->last_ms[2], ctx->ev[2], ctx->ev[3]));
        ctx->last_ms[3] = 0.f;
    }
    return UGSDR_OK;
}

}  // namespace

}  // namespace ugsdr_hip

using namespace ugsdr_hip;

extern "C" {

int ugsdr_hip_stft(ugsdr_hip_ctx* ctx, const float* in_c32, int64_t n, int window, float* out_c32) {
    if (!ctx) {
        set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return stft_impl(ctx, in_c32, n, window, out_c32);
}

int ugsdr_hip_istft(ugsdr_hip_ctx* ctx, const float* cols_c32, int64_t n_cols, float* out_c32) {
    if (!ctx) {
        set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return istft_impl(ctx, cols_c32, n_cols, out_c32);
}

int ugsdr_hip_jse_process(ugsdr_hip_ctx* ctx, float* io_c32, int64_t n, double fs, int window, int on_device,
                          ugsdr_jammer* out, int32_t max_out, int32_t* n_out) {
    if (!ctx) {
        set_error("null context");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    return jse_impl(ctx, io_c32, n, fs, window, on_device, out, max_out, n_out);
}

}  // extern "C"
