// Host-side staging of one acquisition request: the resolved plan, the device
// buffers a request holds while in flight, and the enqueue steps shared by the
// synchronous entry point and the request server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ctx.hpp"
#include "plan.hpp"

namespace ugsdr_hip {

struct AcqPlan {
    int64_t N = 0, K = 0, L = 0, D = 0, P = 0;
    // periodic fold: each of the K blocks is Tsub whole code periods of N
    // samples (replicas of Tsub periods folded to one), searched as an
    // N-length transform of the periods' coherently combined spectra
    int64_t Tsub = 1;
    // upsampled-code search: every replica is one code period of N / ucode
    // chips held for ucode samples each, so the correlation runs as ucode
    // inverse transforms of N / ucode points per row on the PRN-independent
    // spectra Z of each (Doppler bin, block) (0: off)
    int ucode = 0;
    const float2* boxcar = nullptr;   // conj of the ucode-sample boxcar's N-point spectrum
    float zbound = 0.f;               // max over k' of sum_u |H[k' + u N / ucode]|
    int F = 0;
    int fold_mode = 0;        // 0: |.| then fold, 1: fold complex then |.|
    int reshape = 1;
    std::vector<double> cls_frac;
    std::vector<int> cls, shift;
    std::vector<double> doppler;
    std::vector<const float2*> specs;
    // front end
    int fmt = 0;
    int64_t up = 1, down = 1, need = 0, n_in_used = 0;
    int spm_in = 0, epoch_bytes = 0;  // recorded-format geometry (BBP_DDC epochs)
    int64_t in_bytes = 0;             // bytes of the n_in_used input samples
    int apply_tone = 0;
    double f_turns = 0, ph_turns = 0;
    // result scaling
    double ratio = 1, intermediate_frequency = 0, threshold = 3.3;
    // transform plan
    bool small = true;
    int threads = 0;          // correlation workgroup size (0: default)
    RadixPlan plan{};
    FourStep fs{};
    // table layout (bytes) in the request's table buffer
    size_t off_frac = 0, off_dop = 0, off_cls = 0, off_shift = 0, off_specs = 0, table_bytes = 0;
    // upsampled-code search: the correlation's view of Z (bin d's spectra at
    // block k are "class" d * K with no rotation)
    size_t off_ucls = 0, off_ushift = 0;
    // device bytes the request holds while in flight (admission)
    int64_t footprint(bool surfaces) const;
    // the forward transforms of the class spectra run as split four-step
    // launches through a scratch of K x Tsub x F transforms: the four-step
    // lengths, and 16368 = 48 x 341
    bool fwd_scratch_needed() const { return !small || N == 16368; }
    // batch key: requests sharing it run in one correlation launch
    int64_t bucket() const {
        return (N << 6) | ((int64_t)precision << 3) | (ucode ? 4 : 0) | (fold_mode << 1) | (L == N ? 1 : 0);
    }
    // reduced precision: the request's storage type and its replicas' scale
    int precision = UGSDR_PREC_FP32;
    float spec_scale = 1.0f;
};

// Validates the request, resolves replica keys (caller holds ctx->mu).
// allow_ucode: the upsampled-code search may be chosen (it allocates the
// context's boxcar table on first use of a length).
int make_acq_plan(ugsdr_hip_ctx* ctx, const ugsdr_acq_request* r, AcqPlan& pl, bool allow_ucode = true);
// Uploads tables and IQ and runs the front end (unpack, mix, resample).
// Upload only queues the copies and sizes the buffers; Launch only runs the
// kernel (for graph capture, where the copies stay outside the graph).
enum class FrontStep { All, Upload, Launch };
int enqueue_front(const AcqPlan& pl, const ugsdr_acq_request* r, AcqSlot& s, hipStream_t st, bool keep_surf,
                  FrontStep step = FrontStep::All);
int launch_front(const AcqPlan& pl, const ugsdr_acq_request* r, AcqSlot& s, hipStream_t st);
// Forward FFT of every block for every fractional Doppler class.
int enqueue_classes(const AcqPlan& pl, AcqSlot& s, hipStream_t st);
// One correlation launch over every request of the batch (same bucket) with
// `threads` per workgroup (0: default for the length).
int enqueue_corr(const std::vector<const AcqPlan*>& plans, const std::vector<AcqSlot*>& slots, hipStream_t st,
                 DevBuf& scratch, int cu_count, int threads = 0);
// scratch the correlation of this bucket may use (upper bound, any variant)
size_t corr_scratch_bytes(const AcqPlan& p0, int cu_count);
// Grows the slot buffers the class, reduction and selection steps of a search
// use to the sizes those steps reserve, so that they allocate nothing (a
// stream capture must not allocate).  enqueue_front's Upload step sizes the
// rest.
int presize_slot(const AcqPlan& pl, AcqSlot& s, bool peaks);
// workgroup sizes the correlation kernel is built with for a length
std::vector<int> thread_candidates(int64_t N);
int default_threads(int64_t N);
// times every candidate on a synthetic batch of this length and records the
// fastest in ctx->tuned (caller holds ctx->mu); ms_out gets each candidate's time
int autotune_corr(ugsdr_hip_ctx* ctx, int64_t N, int* best, std::vector<float>* ms_out);
// tuned workgroup size for a length (autotunes on first use when enabled)
int tuned_threads(ugsdr_hip_ctx* ctx, int64_t N);
// Per-PRN selection into s.results ([P] ugsdr_acq_result then [P][L] peaks).
// Three-stage correlation for N = 1023*M (M = 4, 8, 12, 16): its workgroup
// size for N (0 if N is not such a length), whether it handles the surface
// length L, and the launch.
int corr1023_threads(int64_t N);
bool corr1023_supports(int64_t N, int64_t L);
// Whole rounds of rows (c.full_rows) run one row per workgroup; the rest are
// split stream-K over the blocks of every workgroup (c.units tail units, the
// ReqDesc tail fields) and merged by a second kernel.  corr1023_grid: the
// workgroups the device holds at once for length N.
struct CorrArgs;
int corr1023_grid(int64_t N, int cus, bool upsampled_code = false);
int launch_corr1023(const CorrArgs& c, int cus, DevBuf& scratch, hipStream_t st);
// upsampled-code search: Z of every (Doppler bin, block) of a request from its
// class spectra (pl.ucode > 0; s.zbuf sized D K N complex)
int launch_zprep(const AcqPlan& pl, AcqSlot& s, hipStream_t st);
// upsampled-code search with codes of 1023 M chips: U phase transforms of
// 1023 M points per row on the three-stage engine (c.ucode = U, c.N = U 1023 M)
int launch_corr1023p(const CorrArgs& c, int cus, hipStream_t st);
// Reduced precision (16-bit storage, matrix-core stages): whether a length
// pair has a plan, the 16-bit packing of the class spectra, and one
// correlation launch over a batch of requests of one bucket.
bool reduced_supported(int64_t N, int64_t L);
// a reduced-precision plan this build runs (upsampled-code or four-step form)
bool reduced_plan_ok(const AcqPlan& pl);
// ProcessGalileo's choice among a satellite's code hypotheses
// (fse.hpp:382-388): per run of h consecutive results, the first of largest
// level, detected when its SNR clears the threshold.  in: n results, out:
// n / h results (h > 1).
void reduce_hypotheses(const ugsdr_acq_result* in, int64_t n, int h, double threshold, ugsdr_acq_result* out);
size_t reduced_scratch_bytes(int64_t N, int max_k, int grid);
int enqueue_reduce_classes(const AcqPlan& pl, AcqSlot& s, hipStream_t st);
// per (block, class) max |X| of the FP32 class spectra into s.block_scale
// (the reduced form's block scales)
int enqueue_block_max(const AcqPlan& pl, AcqSlot& s, hipStream_t st);
int enqueue_corr_reduced(const std::vector<const AcqPlan*>& plans, const std::vector<AcqSlot*>& slots,
                         hipStream_t st, DevBuf& scratch, int cu_count, int threads);
int enqueue_final(const AcqPlan& pl, AcqSlot& s, hipStream_t st, bool peaks);
int device_cu_count(int device);

}  // namespace ugsdr_hip
