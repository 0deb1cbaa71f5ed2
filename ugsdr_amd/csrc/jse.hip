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

// Per column with DETECT: kept[c] = {count, bins of the candidates whose J/N
// is not under 6, in extraction order, 7 spare bytes}; jn[r * ncol + c] =
// J/N of the r-th kept candidate (round-major, so the host reads only the
// rounds it needs).
template <int DETECT>
__global__ __launch_bounds__(256) void stft_kernel(const float2* x, int64_t ncol, const float* window,
                                                   float2* cols, uint8_t* kept, double* jn) {
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
    if (!DETECT) return;
    // IntermediateDetection (jamming_detection.hpp:115-137) on lanes 0..7 of
    // the group (lanes 8..15 mirror them and write nothing)
    float mag[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) mag[k2] = sqrtf(v[k2].x * v[k2].x + v[k2].y * v[k2].y);
    int cbin[kCand];
    float cval[kCand];
#pragma unroll
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
        cbin[r] = bi;
        cval[r] = best;
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
    if (owner && k1 == 0) {
        // J/N = value / sigma; candidates under 6 are dropped in order
        // (jamming_detection.hpp:130-136)
        const double sd = sqrt(q / (kWin - 1));
        uint8_t rec[16] = {};
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < kCand; ++r) {
            const double j = (double)cval[r] / sd;
            if (!(j < 6.0)) {
                rec[1 + cnt] = (uint8_t)cbin[r];
                jn[(int64_t)cnt * ncol + cc] = j;
                ++cnt;
            }
        }
        rec[0] = (uint8_t)cnt;
        *reinterpret_cast<uint4*>(kept + cc * 16) = *reinterpret_cast<const uint4*>(rec);
    }
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

// AnalyzeStftData (jamming_detection.hpp:139-190) over the kept candidates:
// kept[16 c] = count, kept[16 c + 1 + r] = bin of the r-th; jn[r ncol + c].
// Its frequency-ordered map becomes a table over the 128 bins (frequency =
// bin fs / 128 rises with the bin), its sums run in column order, and its
// integer mean count and size_t(size * 0.9) wideband rule are kept.
std::vector<ugsdr_jammer> analyze(const uint8_t* kept, const double* jn, int64_t ncol, size_t rounds, double fs) {
    constexpr double kProb = 0.9;
    std::vector<ugsdr_jammer> dst;
    for (size_t i = 0; i < rounds; ++i) {
        double count[kWin] = {}, jsum[kWin] = {};
        double quantity = 0;
        for (int64_t c = 0; c < ncol; ++c) {
            if (kept[(size_t)c * 16] <= i) continue;
            const int b = kept[(size_t)c * 16 + 1 + i];
            count[b] += 1;
            jsum[b] += jn[i * (size_t)ncol + (size_t)c];
            ++quantity;
        }
        size_t size = 0;
        for (int b = 0; b < kWin; ++b) size += count[b] > 0;
        auto entry = [&](int b) {
            return ugsdr_jammer{0, b, b * fs / kWin, jsum[b] / quantity};
        };
        size_t cnt_wideband = 0;
        int wb = -1;
        for (int b = 0; b < kWin; ++b) {
            if (count[b] == 0) continue;
            if (wb < 0) wb = b;  // the map's first entry
            const double p = count[b] / quantity;
            if (p >= kProb) {
                ugsdr_jammer d = entry(b);
                d.type = UGSDR_JAMMER_NARROWBAND;
                dst.push_back(d);
                break;
            }
            if (p >= kProb / (double)size) {
                ++cnt_wideband;
                wb = b;
            }
        }
        if (cnt_wideband >= (size_t)(size * kProb)) {
            ugsdr_jammer d = entry(wb);
            d.type = UGSDR_JAMMER_WIDEBAND;
            dst.push_back(d);
        }
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
                       nullptr, nullptr);
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
    const size_t b_kept = (size_t)ncol * 16, o_jn = (b_kept + 255) / 256 * 256;
    UGSDR_HIP_TRY(ctx->results.reserve(o_jn + (size_t)ncol * kCand * 8));
    uint8_t* d_kept = ctx->results.as<uint8_t>();
    double* d_jn = reinterpret_cast<double*>(ctx->results.as<char>() + o_jn);
    const unsigned blocks = (unsigned)((ncol + kColsPerBlock - 1) / kColsPerBlock);
    ctx->mark(0);
    hipLaunchKernelGGL(stft_kernel<1>, dim3(blocks), dim3(256), 0, st, x, ncol, ctx->tables.as<float>(),
                       ctx->sig.as<float2>(), d_kept, d_jn);
    UGSDR_HIP_TRY(hipGetLastError());
    ctx->mark(1);
    // the kept candidates' bins first; their J/N only for the rounds the
    // mean count asks for
    std::vector<uint8_t> kept(b_kept);
    UGSDR_HIP_TRY(hipMemcpyAsync(kept.data(), d_kept, b_kept, hipMemcpyDeviceToHost, st));
    UGSDR_HIP_TRY(hipStreamSynchronize(st));
    size_t total = 0;
    for (int64_t c = 0; c < ncol; ++c) total += kept[(size_t)c * 16];
    const size_t rounds = total / (size_t)ncol;
    std::vector<double> jn(rounds * (size_t)ncol);
    if (rounds) {
        UGSDR_HIP_TRY(hipMemcpyAsync(jn.data(), d_jn, jn.size() * 8, hipMemcpyDeviceToHost, st));
        UGSDR_HIP_TRY(hipStreamSynchronize(st));
    }
    const std::vector<ugsdr_jammer> det = analyze(kept.data(), jn.data(), ncol, rounds, fs);
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
