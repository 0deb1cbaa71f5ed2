// Correlation batch descriptors and the workgroup reductions that turn a
// correlation surface into row statistics, shared by the correlation kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_fft.hpp"
#include "plan.hpp"

namespace ugsdr_hip {

using namespace ugsdr_dev;

// One search request inside a correlation batch.  Every request of a batch
// shares N, the surface length and the fold mode (one kernel instantiation);
// K, the Doppler bins and the PRN set are per request.
struct ReqDesc {
    const float2* X;             // [K][F][N] class spectra
    const float2* const* specs;  // [P] device pointers to conj replica spectra
    const int* cls;              // [D] class of each Doppler bin
    const int* shift;            // [D] integer bin rotation of each Doppler bin
    RowStat* rowstats;           // [P*D]
    float* surfaces;             // optional [P*D][L]
    int K, F, D, P;
    int row_start;               // first batch row of this request
    // three-stage kernel: rows from tail_row0 on (request-local) belong to
    // the batch tail, whose (row, block) units start at tail_unit_start
    int tail_row0;
    int tail_unit_start;
    // reduced precision: X holds 16-bit pairs and block_inv[k*F + c] undoes
    // the power-of-two scale of class block (k, c); spec_inv undoes the
    // replica scale
    const float* block_inv;
    float spec_inv;
    int pad2;
};

// requests one correlation launch carries (their descriptors travel in the
// kernel arguments, read through the scalar cache; no upload)
constexpr int kMaxBatch = 16;
// periodic-fold searches combine at most this many code periods per block
constexpr int kMaxPeriods = 64;

struct CorrArgs {
    RadixPlan plan;
    FourStep fs;
    int N, L;
    int fold_mode;          // 0: |.| then fold, 1: fold complex then |.|
    int rows;               // total rows (sum of P*D)
    int n_req;
    ReqDesc reqs[kMaxBatch];  // [n_req], ascending row_start
    float2* scratch;        // large path: per workgroup 2N complex + L floats;
                            // stream-K path: [grid][2][L] partial surfaces
    float inv_n;
    int full_rows;          // three-stage kernel: rows run one per workgroup
    int units;              // (row, block) units of the tail after them
    int max_k;              // reduced path: most blocks of any request
    int sk_grid;            // three-stage kernel: workgroups of the tail split
    int n_merge;            // three-stage kernel: tail rows to merge
    int ucode;              // upsampled-code search: samples per chip (0: off)
    // upsampled-code search: rows are ordered Doppler-bin-major, and the
    // workgroups the device holds at once (row_group, a multiple of the 8
    // XCDs) take them so that each XCD's share is consecutive rows, which
    // read the same spectra Z through that XCD's L2 (0: no reordering)
    int row_group;
};

// upsampled-code search: the row a main-kernel workgroup of index b takes
// (workgroups are dispatched round-robin over the 8 XCDs)
__device__ __forceinline__ int up_row_of(int b, int full_rows, int G) {
    if (G <= 0 || b >= full_rows / G * G) return b;
    const int g = b / G, w = b - g * G;
    return g * G + (w & 7) * (G >> 3) + (w >> 3);
}

// a request-local row's (PRN, Doppler bin) and its result slot (PRN-major):
// Doppler-bin-major rows for the upsampled-code search, PRN-major otherwise
struct RowRef {
    int p, di, slot;
};
__device__ __forceinline__ RowRef row_ref(const ReqDesc& q, int local, bool bin_major) {
    RowRef r;
    if (bin_major) {
        r.di = local / q.P;
        r.p = local - r.di * q.P;
    } else {
        r.p = local / q.D;
        r.di = local - r.p * q.D;
    }
    r.slot = r.p * q.D + r.di;
    return r;
}

// the request owning a batch row (uniform per workgroup: kept in scalar registers)
__device__ __forceinline__ int find_req(const CorrArgs& a, int row) {
    int r = 0;
    while (r + 1 < a.n_req && a.reqs[r + 1].row_start <= row) ++r;
    return __builtin_amdgcn_readfirstlane(r);
}

// the request owning tail unit u (requests before the tail have no tail
// units and the same start, so the last request at or below u owns it)
__device__ __forceinline__ int find_req_unit(const CorrArgs& a, int u) {
    int r = 0;
    while (r + 1 < a.n_req && a.reqs[r + 1].tail_unit_start <= u) ++r;
    return __builtin_amdgcn_readfirstlane(r);
}

// Stream-K partition: workgroup w of G owns units [begin(w), begin(w + 1)).
__host__ __device__ inline int sk_begin(int units, int G, int w) { return (int)((int64_t)units * w / G); }
// the workgroup owning unit x
__host__ __device__ inline int sk_owner(int units, int G, int x) {
    return (int)(((int64_t)(x + 1) * G - 1) / units);
}

// block-wide reductions over T threads (T a multiple of 64: T/64 full waves)
template <int T>
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
    v = wave_sum_d(v);
    const int w = tid >> 6, l = tid & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < T / 64; ++i) s += red[i];
    return s;
}

// top-2 with first-index tie break: (v1,i1) the max, v2 the runner-up value
__device__ __forceinline__ void top2_merge(float& v1, int& i1, float& v2, float ov1, int oi1, float ov2) {
    if (ov1 > v1 || (ov1 == v1 && oi1 < i1)) {
        v2 = fmaxf(v1, ov2);
        v1 = ov1;
        i1 = oi1;
    } else {
        v2 = fmaxf(v2, ov1);
    }
}

template <int T>
__device__ __forceinline__ void block_top2(float& v1, int& i1, float& v2, double* red, int tid) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float a = __shfl_xor(v1, off, 64);
        const int b = __shfl_xor(i1, off, 64);
        const float c = __shfl_xor(v2, off, 64);
        top2_merge(v1, i1, v2, a, b, c);
    }
    float* rf = reinterpret_cast<float*>(red);
    int* ri = reinterpret_cast<int*>(red);
    const int w = tid >> 6, l = tid & 63;
    __syncthreads();
    if (l == 0) {
        rf[3 * w] = v1;
        ri[3 * w + 1] = i1;
        rf[3 * w + 2] = v2;
    }
    __syncthreads();
    v1 = rf[0];
    i1 = ri[1];
    v2 = rf[2];
#pragma unroll
    for (int i = 1; i < T / 64; ++i) top2_merge(v1, i1, v2, rf[3 * i], ri[3 * i + 1], rf[3 * i + 2]);
}

// Row statistics: first-index max, runner-up, mean and n-1 sigma (two-pass).
template <int T>
__device__ __forceinline__ void stats_write(float v1, int i1, float v2, double mean, double q, int L, int tid,
                                            RowStat* out, float inv_n) {
    if (tid == 0) {
        RowStat r;
        r.level = v1 * inv_n;
        r.index = i1;
        r.second = v2 > 0.f ? v2 * inv_n : 0.f;
        r.mean = L >= 2 ? (float)(mean * inv_n) : 0.f;
        r.stddev = L >= 2 ? (float)(sqrt(q / (double)(L - 1)) * inv_n) : 0.f;
        *out = r;
    }
}

// surface held in registers: S[i] is element j = tid + i*T (j < L)
template <int T, int SR>
__device__ __forceinline__ void row_stats_reg(const float (&S)[SR], int L, double* red, int tid, RowStat* out,
                                              float inv_n) {
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
        const int j = tid + i * T;
        if (j < L) {
            top2_merge(v1, i1, v2, S[i], j, -1.0f);
            sum += (double)S[i];
        }
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double q = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
        const int j = tid + i * T;
        if (j < L) {
            const double d = (double)S[i] - mean;
            q += d * d;
        }
    }
    q = block_sum<T>(q, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, q, L, tid, out, inv_n);
}

// surface in (workgroup-private) global memory
template <int T>
__device__ __forceinline__ void row_stats_glob(const float* S, int L, double* red, int tid, RowStat* out,
                                               float inv_n) {
    float v1 = -1.0f, v2 = -1.0f;
    int i1 = 0x7fffffff;
    double sum = 0.0;
    for (int j = tid; j < L; j += T) {
        const float s = S[j];
        top2_merge(v1, i1, v2, s, j, -1.0f);
        sum += (double)s;
    }
    block_top2<T>(v1, i1, v2, red, tid);
    sum = block_sum<T>(sum, red + 64, tid);
    const double mean = sum / (double)L;
    double q = 0.0;
    for (int j = tid; j < L; j += T) {
        const double d = (double)S[j] - mean;
        q += d * d;
    }
    q = block_sum<T>(q, red + 64, tid);
    stats_write<T>(v1, i1, v2, mean, q, L, tid, out, inv_n);
}

}  // namespace ugsdr_hip
