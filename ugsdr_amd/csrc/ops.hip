// Per-operator drop-ins behind the Hip* CRTP classes: one call = one reference
// operator call, host data in and out.  This is the slow, general path (any
// length); the fused acquisition in acq.hip is the fast path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "dev_fft.hpp"
#include "formats.hpp"
#include "internal.hpp"

using namespace ugsdr_dev;

namespace ugsdr_hip {

static inline int grid_for(int64_t n, int block = 256) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 8192));
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// x[n] *= exp(j 2 pi (n f/fs + phase/2pi)), phase in FP64 cycles
__global__ void mix_kernel(const float2* in, float2* out, int64_t n, double f_turns, double ph_turns) {
    GRID_STRIDE(i, n) {
        double t = fma((double)i, f_turns, ph_turns);
        t -= rint(t);
        float s, c;
        sincosf((float)(2.0 * kPi * t), &s, &c);
        out[i] = cmul(in[i], make_float2(c, s));
    }
}

// TableMixer's 80-entry sine table (src/mixer/table_mixer.hpp:31-47), built on
// the host in FP64 and rounded once to float, so that the device uses the very
// values the reference's host build computes.
struct SineTable80 {
    float v[80];
};

// TableMixer::Process (table_mixer.hpp:78-101) in place: sample i takes the
// top 6 bits of the 32-bit accumulator holder0 + i * adder (nco.hpp:17-20,
// wrapping), cos at entry + 16 and sin at the entry; the complex product is
// formed from separately rounded products, as the reference's
// complex<float> multiply is (no fused multiply-add).
__global__ void table_mix_kernel(float2* io, int64_t n, uint32_t holder0, uint32_t adder, SineTable80 tab) {
#pragma clang fp contract(off)
    __shared__ float t[80];
    if (threadIdx.x < 80) t[threadIdx.x] = tab.v[threadIdx.x];
    __syncthreads();
    GRID_STRIDE(i, n) {
        const uint32_t h = holder0 + (uint32_t)(uint64_t)i * adder;
        const uint32_t idx = h >> 26;
        const float c = t[idx + 16], s = t[idx];
        const float2 x = io[i];
        const float ac = x.x * c, bs = x.y * s, as = x.x * s, bc = x.y * c;
        io[i] = make_float2(ac - bs, as + bc);
    }
}

// One generic Stockham stage, any radix R, one thread per output element:
// out[(j/Ns)*Ns*R + j%Ns + q*Ns] = sum_r in[j + r*N/R] exp(-2 pi i r (j%Ns + q Ns)/(Ns R))
__global__ void dft_stage_kernel(const float2* in, float2* out, int64_t N, int64_t R, int64_t Ns) {
    const int64_t stride = N / R;
    GRID_STRIDE(o, N) {
        const int64_t q = o / stride;
        const int64_t j = o - q * stride;
        const int64_t k = j % Ns;
        const int64_t span = Ns * R;
        const int64_t step = (k + q * Ns) % span;
        double ar = 0.0, ai = 0.0;
        int64_t m = 0;  // r * step mod span
        for (int64_t r = 0; r < R; ++r) {
            const float2 x = in[j + r * stride];
            double s, c;
            sincospi(-2.0 * (double)m / (double)span, &s, &c);
            ar += (double)x.x * c - (double)x.y * s;
            ai += (double)x.x * s + (double)x.y * c;
            m += step;
            if (m >= span) m -= span;
        }
        out[(j - k) * R + k + q * Ns] = make_float2((float)ar, (float)ai);
    }
}

__global__ void conj_scale_kernel(float2* x, int64_t n, float scale) {
    GRID_STRIDE(i, n) x[i] = make_float2(x[i].x * scale, -x[i].y * scale);
}
__global__ void real_to_complex_kernel(const float* in, float2* out, int64_t n) {
    GRID_STRIDE(i, n) out[i] = make_float2(in[i], 0.f);
}
__global__ void cmul_kernel(float2* x, const float2* y, int64_t n) {
    GRID_STRIDE(i, n) x[i] = cmul(x[i], y[i]);
}
__global__ void abs_kernel(const float2* in, float* out, int64_t n) {
    GRID_STRIDE(i, n) out[i] = sqrtf(fmaf(in[i].x, in[i].x, in[i].y * in[i].y));
}
// dst[j] = ((src[j] + src[j+b]) + src[j+2b]) + ...
__global__ void reshape_sum_kernel(const float* in, float* out, int64_t n, int64_t block) {
    GRID_STRIDE(j, block) {
        float acc = in[j];
        for (int64_t b = block; b < n; b += block) acc += in[b + j];
        out[j] = acc;
    }
}
__global__ void reshape_sum_c_kernel(const float2* in, float2* out, int64_t n, int64_t block) {
    GRID_STRIDE(j, block) {
        float2 acc = in[j];
        for (int64_t b = block; b < n; b += block) acc = cadd(acc, in[b + j]);
        out[j] = acc;
    }
}
__global__ void upsample_kernel(const float* in, int64_t len, float* out, int64_t samples) {
    GRID_STRIDE(i, samples) out[i] = in[(i * len / samples) % len];
}
// sample-and-hold to lcm then boxcar down (zero-padded tail)
__global__ void resample_kernel(const float2* in, int64_t n_in, float2* out, int64_t n_out, int64_t up,
                                int64_t down) {
    GRID_STRIDE(o, n_out) {
        double sr = 0.0, si = 0.0;
        const int64_t u0 = o * down;
        for (int64_t u = u0; u < u0 + down; ++u) {
            const int64_t i = u / up;
            if (i >= n_in) break;
            sr += in[i].x;
            si += in[i].y;
        }
        out[o] = make_float2((float)(sr / (double)down), (float)(si / (double)down));
    }
}

constexpr int kRedT = 1024;

__device__ __forceinline__ double block_reduce_sum(double v, double* sh) {
    v = wave_sum_d(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    return s;
}

// single workgroup: first index of the max
__global__ void max_index_kernel(const float* in, int64_t n, float* value, int64_t* index) {
    __shared__ float sv[16];
    __shared__ long long si[16];
    float v = -INFINITY;
    long long idx = 0x7fffffffffffffffll;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        if (in[i] > v) { v = in[i]; idx = i; }
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(v, off, 64);
        const long long oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = idx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            if (sv[i] > v || (sv[i] == v && si[i] < idx)) { v = sv[i]; idx = si[i]; }
        *value = v;
        *index = idx;
    }
}

// single workgroup: mean and sample sigma, two-pass in FP64
__global__ void mean_std_kernel(const float* in, int64_t n, float* out) {
    __shared__ double sh[16];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += in[i];
    s = block_reduce_sum(s, sh);
    const double mean = s / (double)n;
    double q = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = (double)in[i] - mean;
        q += d * d;
    }
    q = block_reduce_sum(q, sh);
    if (threadIdx.x == 0) {
        out[0] = (float)mean;
        out[1] = (float)std::sqrt(q / (double)(n - 1));
    }
}

// single workgroup: complex x real dot product, FP64 accumulation
__global__ void dot_kernel(const float2* s, const float* c, int64_t n, float* out) {
    __shared__ double sh[16];
    double re = 0.0, im = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        re += (double)s[i].x * c[i];
        im += (double)s[i].y * c[i];
    }
    re = block_reduce_sum(re, sh);
    im = block_reduce_sum(im, sh);
    if (threadIdx.x == 0) {
        out[0] = (float)re;
        out[1] = (float)im;
    }
}

// ---------------------------------------------------------------- host side
struct Lock {
    ugsdr_hip_ctx* ctx;
    std::lock_guard<std::mutex> g;
    explicit Lock(ugsdr_hip_ctx* c) : ctx(c), g(c->mu) {}
};

static int h2d(ugsdr_hip_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
    UGSDR_HIP_TRY(b.reserve(bytes));
    if (bytes) UGSDR_HIP_TRY(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return UGSDR_OK;
}
static int d2h(ugsdr_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes) UGSDR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

static void factor(int64_t n, std::vector<int64_t>& f) {
    f.clear();
    int64_t m = n;
    while (m % 4 == 0) { f.push_back(4); m /= 4; }
    while (m % 2 == 0) { f.push_back(2); m /= 2; }
    for (int64_t p = 3; p * p <= m; p += 2)
        while (m % p == 0) { f.push_back(p); m /= p; }
    if (m > 1) f.push_back(m);
}

// forward DFT of d (n complex, device) in place using tmp (n complex)
int dft_device(ugsdr_hip_ctx* ctx, float2* d, float2* tmp, int64_t n, int inverse) {
    if (n <= 1) return UGSDR_OK;
    std::vector<int64_t> f;
    factor(n, f);
    // each stage costs n * radix: a length with a large prime factor would
    // make one stage quadratic in that factor
    if (f.back() > 4096) {
        set_error("DFT length " + std::to_string(n) + " has the prime factor " + std::to_string(f.back()) +
                  " (at most 4096 supported)");
        return UGSDR_ERR_UNSUPPORTED;
    }
    if (inverse) hipLaunchKernelGGL(conj_scale_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, n, 1.0f);
    float2* a = d;
    float2* b = tmp;
    int64_t Ns = 1;
    for (int64_t R : f) {
        hipLaunchKernelGGL(dft_stage_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, a, b, n, R, Ns);
        Ns *= R;
        std::swap(a, b);
    }
    if (a != d) UGSDR_HIP_TRY(hipMemcpyAsync(d, a, n * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream));
    if (inverse)
        hipLaunchKernelGGL(conj_scale_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, n,
                           (float)(1.0 / (double)n));
    UGSDR_HIP_TRY(hipGetLastError());
    return UGSDR_OK;
}

static int mix_common(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n, double fs, double f, double phase,
                      bool inplace_semantics) {
    UGSDR_CHECK_ARG(ctx && in && out && fs > 0 && n >= 0, "bad mixer arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    phase = std::fmod(phase, 2 * M_PI);  // Mixer::FixPhase (mixer.hpp:20-25)
    if (phase < 0) phase += 2 * M_PI;
    if (inplace_semantics && std::fabs(f) == 0.0) {
        if (in != out) std::copy(in, in + 2 * n, out);
        return UGSDR_OK;
    }
    if (f < 0) f = fs + f;
    if (int rc = h2d(ctx, ctx->aux, in, n * 8)) return rc;
    hipLaunchKernelGGL(mix_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->aux.as<float2>(),
                       ctx->aux.as<float2>(), n, f / fs, phase / (2 * M_PI));
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux.ptr, n * 8);
}

// Mixer<TableMixer>::Translate, in-place overload (mixer.hpp:33-41): FixPhase,
// f == 0 returns early, f < 0 folds to fs + f, then TableMixer::Process.
static int table_mix_common(ugsdr_hip_ctx* ctx, float* io, int64_t n, double fs, double f, double phase,
                            int on_device) {
    UGSDR_CHECK_ARG(ctx && io && fs > 0 && n >= 0, "bad table mixer arguments");
    phase = std::fmod(phase, 2 * M_PI);
    if (phase < 0) phase += 2 * M_PI;
    if (std::fabs(f) == 0.0) return UGSDR_OK;
    if (f < 0) f = fs + f;
    UGSDR_CHECK_ARG(f >= 0, "table mixer frequency below -fs");
    SineTable80 tab;
    const double step = 5 * M_PI / 80.0 / 2;   // table_mixer.hpp:37
    for (int i = 0; i < 80; ++i) tab.v[i] = (float)(1.0 * std::sin((double)i * step));
    // NumericallyControlledOscillator<32> (nco.hpp:17-20): truncating casts,
    // the 32-bit holder keeps the low 32 bits
    const uint32_t adder = (uint32_t)(uint64_t)(4294967296.0 * f / fs);
    const uint32_t holder = (uint32_t)(uint64_t)(4294967296.0 * phase / (2 * M_PI));
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    float2* d = reinterpret_cast<float2*>(io);
    if (!on_device) {
        if (int rc = h2d(ctx, ctx->aux, io, n * 8)) return rc;
        d = ctx->aux.as<float2>();
    }
    if (n > 0)
        hipLaunchKernelGGL(table_mix_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, n, holder, adder, tab);
    UGSDR_HIP_TRY(hipGetLastError());
    if (!on_device) return d2h(ctx, io, ctx->aux.ptr, n * 8);
    UGSDR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return UGSDR_OK;
}

// ---------------------------------------------------------------- formats
__global__ void unpack_kernel(SampleSource src, float2* out, int64_t n) {
    GRID_STRIDE(i, n) out[i] = load_sample(src, i);
}

int format_geometry(int fmt, double fs, int& spm, int& epoch_bytes) {
    spm = 0;
    epoch_bytes = 0;
    if (fmt < UGSDR_FMT_CI8 || fmt > UGSDR_FMT_NT1065_4) {
        set_error("unknown sample format");
        return UGSDR_ERR_ARG;
    }
    if (fmt != UGSDR_FMT_BBP_DDC) return UGSDR_OK;
    const double per_ms = fs / 1e3;
    if (!(per_ms >= 2 && per_ms < (1 << 26)) || per_ms != std::floor(per_ms)) {
        set_error("BBP_DDC needs a sampling rate of a whole number of kHz");
        return UGSDR_ERR_ARG;
    }
    spm = (int)per_ms;
    // words = spm * 4 / 64 + 1 (two 2-bit components a sample, one header
    // word), rounded up to an even count (signal_parameters.hpp:74-80)
    epoch_bytes = 16 * ((spm + 16 + 31) / 32);
    return UGSDR_OK;
}

int64_t format_bytes(int fmt, int64_t n, int spm, int epoch_bytes) {
    switch (fmt) {
        case UGSDR_FMT_CI8: return 2 * n;
        case UGSDR_FMT_CI16: return 4 * n;
        case UGSDR_FMT_CF32: return 8 * n;
        case UGSDR_FMT_BBP_DDC: return (n + spm - 1) / spm * (int64_t)epoch_bytes;
        default: return n;  // R8, NT1065_*
    }
}

int verify_bbp_headers(const void* host, int64_t n_epochs, int epoch_bytes) {
    const unsigned char* b = static_cast<const unsigned char*>(host);
    auto counter = [&](int64_t e) {
        uint64_t w;
        memcpy(&w, b + e * epoch_bytes, 8);
        return w & ((1u << 24) - 1);
    };
    for (int64_t e = 1; e < n_epochs; ++e)
        if (counter(e) - counter(e - 1) != 1) {
            set_error("Header ms count mismatch");  // signal_parameters.hpp:41-42
            return UGSDR_ERR_ARG;
        }
    return UGSDR_OK;
}

}  // namespace ugsdr_hip

using namespace ugsdr_hip;

extern "C" {

int ugsdr_hip_mix(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n, double fs, double f, double phase) {
    return mix_common(ctx, in, out, n, fs, f, phase, false);
}

int ugsdr_hip_mix_inplace(ugsdr_hip_ctx* ctx, float* io, int64_t n, double fs, double f, double phase) {
    return mix_common(ctx, io, io, n, fs, f, phase, true);
}

int ugsdr_hip_table_mix(ugsdr_hip_ctx* ctx, float* io, int64_t n, double fs, double f, double phase,
                        int on_device) {
    return table_mix_common(ctx, io, n, fs, f, phase, on_device);
}

int ugsdr_hip_dft(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n, int inverse) {
    UGSDR_CHECK_ARG(ctx && in && out && n >= 0, "bad dft arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 8)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(n * 8));
    if (int rc = dft_device(ctx, ctx->aux.as<float2>(), ctx->aux2.as<float2>(), n, inverse)) return rc;
    return d2h(ctx, out, ctx->aux.ptr, n * 8);
}

int ugsdr_hip_mf_prepare(ugsdr_hip_ctx* ctx, const float* code, float* spec, int64_t n) {
    UGSDR_CHECK_ARG(ctx && code && spec && n > 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux2, code, n * 4)) return rc;
    UGSDR_HIP_TRY(ctx->aux.reserve(n * 8));
    UGSDR_HIP_TRY(ctx->sig.reserve(n * 8));
    hipLaunchKernelGGL(real_to_complex_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->aux2.as<float>(),
                       ctx->aux.as<float2>(), n);
    if (int rc = dft_device(ctx, ctx->aux.as<float2>(), ctx->sig.as<float2>(), n, 0)) return rc;
    hipLaunchKernelGGL(conj_scale_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->aux.as<float2>(), n, 1.0f);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, spec, ctx->aux.ptr, n * 8);
}

static int mf_filter_dev(ugsdr_hip_ctx* ctx, int64_t n) {
    // ctx->aux: signal, ctx->aux2: spectrum, ctx->sig: temp
    UGSDR_HIP_TRY(ctx->sig.reserve(n * 8));
    if (int rc = dft_device(ctx, ctx->aux.as<float2>(), ctx->sig.as<float2>(), n, 0)) return rc;
    hipLaunchKernelGGL(cmul_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->aux.as<float2>(),
                       ctx->aux2.as<float2>(), n);
    return dft_device(ctx, ctx->aux.as<float2>(), ctx->sig.as<float2>(), n, 1);
}

int ugsdr_hip_mf_filter(ugsdr_hip_ctx* ctx, const float* in, const float* spec, float* out, int64_t n) {
    UGSDR_CHECK_ARG(ctx && in && spec && out && n > 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 8)) return rc;
    if (int rc = h2d(ctx, ctx->aux2, spec, n * 8)) return rc;
    if (int rc = mf_filter_dev(ctx, n)) return rc;
    return d2h(ctx, out, ctx->aux.ptr, n * 8);
}

int ugsdr_hip_mf_filter_code(ugsdr_hip_ctx* ctx, const float* in, const float* code, float* out, int64_t n) {
    UGSDR_CHECK_ARG(ctx && in && code && out && n > 0, "bad arguments");
    std::vector<float> spec(2 * n);
    if (int rc = ugsdr_hip_mf_prepare(ctx, code, spec.data(), n)) return rc;
    return ugsdr_hip_mf_filter(ctx, in, spec.data(), out, n);
}

int ugsdr_hip_abs(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n) {
    UGSDR_CHECK_ARG(ctx && in && out && n >= 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 8)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(n * 4));
    hipLaunchKernelGGL(abs_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->aux.as<float2>(),
                       ctx->aux2.as<float>(), n);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux2.ptr, n * 4);
}

int ugsdr_hip_reshape_sum(ugsdr_hip_ctx* ctx, const float* in, float* out, int64_t n, int64_t block, int complex_) {
    UGSDR_CHECK_ARG(ctx && in && out && n >= 0 && block > 0, "bad arguments");
    if (n % block) {
        set_error("Signal size is not divided by block size");  // reshape_and_sum.hpp:16-17
        return UGSDR_ERR_SIZE;
    }
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    const size_t es = complex_ ? 8 : 4;
    if (int rc = h2d(ctx, ctx->aux, in, n * es)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(block * es));
    if (complex_)
        hipLaunchKernelGGL(reshape_sum_c_kernel, dim3(grid_for(block)), dim3(256), 0, ctx->stream,
                           ctx->aux.as<float2>(), ctx->aux2.as<float2>(), n, block);
    else
        hipLaunchKernelGGL(reshape_sum_kernel, dim3(grid_for(block)), dim3(256), 0, ctx->stream, ctx->aux.as<float>(),
                           ctx->aux2.as<float>(), n, block);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux2.ptr, block * es);
}

int ugsdr_hip_max_index(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* value, int64_t* index) {
    UGSDR_CHECK_ARG(ctx && in && value && index && n > 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 4)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(16));
    hipLaunchKernelGGL(max_index_kernel, dim3(1), dim3(kRedT), 0, ctx->stream, ctx->aux.as<float>(), n,
                       ctx->aux2.as<float>(), reinterpret_cast<int64_t*>(ctx->aux2.as<char>() + 8));
    UGSDR_HIP_TRY(hipGetLastError());
    char buf[16];
    if (int rc = d2h(ctx, buf, ctx->aux2.ptr, 16)) return rc;
    std::memcpy(value, buf, 4);
    std::memcpy(index, buf + 8, 8);
    return UGSDR_OK;
}

int ugsdr_hip_mean_std(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* mean, float* sigma) {
    UGSDR_CHECK_ARG(ctx && in && mean && sigma && n >= 0, "bad arguments");
    if (n < 2) {  // mean_stddev.hpp:33-34
        *mean = 0.f;
        *sigma = 0.f;
        return UGSDR_OK;
    }
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 4)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(8));
    hipLaunchKernelGGL(mean_std_kernel, dim3(1), dim3(kRedT), 0, ctx->stream, ctx->aux.as<float>(), n,
                       ctx->aux2.as<float>());
    UGSDR_HIP_TRY(hipGetLastError());
    float buf[2];
    if (int rc = d2h(ctx, buf, ctx->aux2.ptr, 8)) return rc;
    *mean = buf[0];
    *sigma = buf[1];
    return UGSDR_OK;
}

static int64_t gcd64(int64_t a, int64_t b) {
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}

int64_t ugsdr_resample_len(int64_t n, int64_t new_fs, int64_t old_fs) {
    if (n < 0 || new_fs <= 0 || old_fs <= 0) return -1;
    const int64_t lcm = new_fs / gcd64(new_fs, old_fs) * old_fs;
    const int64_t up_n = lcm / old_fs * n, down = lcm / new_fs;
    return down == 1 ? up_n : (up_n + down - 1) / down;
}

int ugsdr_hip_resample(ugsdr_hip_ctx* ctx, const float* in, int64_t n, int64_t new_fs, int64_t old_fs, float* out) {
    UGSDR_CHECK_ARG(ctx && in && out && n >= 0 && new_fs > 0 && old_fs > 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t lcm = new_fs / gcd64(new_fs, old_fs) * old_fs;
    const int64_t up = lcm / old_fs, down = lcm / new_fs;
    const int64_t n_out = ugsdr_resample_len(n, new_fs, old_fs);
    if (int rc = h2d(ctx, ctx->aux, in, n * 8)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(n_out * 8));
    hipLaunchKernelGGL(resample_kernel, dim3(grid_for(n_out)), dim3(256), 0, ctx->stream, ctx->aux.as<float2>(), n,
                       ctx->aux2.as<float2>(), n_out, up, down);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux2.ptr, n_out * 8);
}

int ugsdr_hip_upsample(ugsdr_hip_ctx* ctx, const float* in, int64_t n, float* out, int64_t samples) {
    UGSDR_CHECK_ARG(ctx && in && out && n > 0 && samples >= 0, "bad arguments");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, n * 4)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(samples * 4));
    hipLaunchKernelGGL(upsample_kernel, dim3(grid_for(samples)), dim3(256), 0, ctx->stream, ctx->aux.as<float>(), n,
                       ctx->aux2.as<float>(), samples);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux2.ptr, samples * 4);
}

int ugsdr_hip_correlate(ugsdr_hip_ctx* ctx, const float* signal, int64_t n_signal, const float* code, int64_t n_code,
                        float* out) {
    UGSDR_CHECK_ARG(ctx && out && n_signal >= 0, "bad arguments");
    if (n_signal != n_code) {
        set_error("Size mismatch");  // correlator.hpp:30-31
        return UGSDR_ERR_SIZE;
    }
    if (n_signal == 0) {
        out[0] = out[1] = 0.f;
        return UGSDR_OK;
    }
    UGSDR_CHECK_ARG(signal && code, "null input");
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, signal, n_signal * 8)) return rc;
    if (int rc = h2d(ctx, ctx->aux2, code, n_code * 4)) return rc;
    UGSDR_HIP_TRY(ctx->rows.reserve(8));
    hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(kRedT), 0, ctx->stream, ctx->aux.as<float2>(), ctx->aux2.as<float>(),
                       n_signal, ctx->rows.as<float>());
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->rows.ptr, 8);
}


int64_t ugsdr_format_bytes(int fmt, int64_t n_samples, double fs) {
    int spm, eb;
    if (int rc = format_geometry(fmt, fs, spm, eb)) return rc;
    if (n_samples < 0) {
        set_error("negative sample count");
        return UGSDR_ERR_ARG;
    }
    return format_bytes(fmt, n_samples, spm, eb);
}

int ugsdr_hip_unpack(ugsdr_hip_ctx* ctx, const void* in, int fmt, int64_t n, double fs, float* out) {
    UGSDR_CHECK_ARG(ctx && in && out && n >= 0, "bad arguments");
    UGSDR_CHECK_ARG(n < ((int64_t)1 << 31), "at most 2^31 - 1 samples a call");
    SampleSource src{};
    if (int rc = format_geometry(fmt, fs, src.spm, src.epoch_bytes)) return rc;
    src.fmt = fmt;
    const int64_t bytes = format_bytes(fmt, n, src.spm, src.epoch_bytes);
    if (fmt == UGSDR_FMT_BBP_DDC)
        if (int rc = verify_bbp_headers(in, bytes / src.epoch_bytes, src.epoch_bytes)) return rc;
    Lock lk(ctx);
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    if (int rc = h2d(ctx, ctx->aux, in, bytes)) return rc;
    UGSDR_HIP_TRY(ctx->aux2.reserve(n * 8));
    src.p = ctx->aux.ptr;
    hipLaunchKernelGGL(unpack_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, src, ctx->aux2.as<float2>(), n);
    UGSDR_HIP_TRY(hipGetLastError());
    return d2h(ctx, out, ctx->aux2.ptr, n * 8);
}

}  // extern "C"
