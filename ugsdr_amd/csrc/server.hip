// Continuous-batching acquisition server.
//
//   submit()   -> validates + plans the request (ctx lock), copies it, queues it
//   scheduler  -> admits queued requests in arrival order while a slot is free
//                 and the in-flight device bytes stay under budget; groups the
//                 admitted ones by transform bucket (N, fold, surface length);
//                 per group: front end + class FFTs per request, ONE correlation
//                 launch for the whole group, per-request finalize and result
//                 copy; each round goes to the next stream of a small pool so
//                 rounds overlap on the device
//   completer  -> polls the rounds' events, stamps latencies, frees slots
//
// With round graphs on, a round's device steps (every request's front end,
// class spectra and reduction, the batched correlation, every request's
// selection) are captured and run as one graph launch; the uploads and result
// copies around them stay outside the graph on the same stream.  Executable graphs
// are cached per stream and topology (the captured sequence of kernels): a
// round whose topology has an idle executable updates that executable's node
// parameters (hipGraphExecUpdate) and replays it.  Otherwise the round's steps
// run eagerly and its capture goes to an instantiation thread, which adds an
// executable for the topology's later rounds (at most kGraphPool per
// topology, shared by the streams), so no round waits for hipGraphInstantiate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "acq.hpp"
#include "formats.hpp"
#include "corr_common.hpp"
#include "ctx.hpp"
#include "internal.hpp"

namespace ugsdr_hip {

using Clock = std::chrono::steady_clock;

struct ServerSlot {
    AcqSlot dev;
    PinnedBuf iq_stage, results;
};

struct Job {
    uint64_t ticket = 0;
    ugsdr_acq_request req{};
    std::vector<double> doppler;
    std::vector<uint64_t> keys;
    std::vector<char> iq;           // host IQ copy (empty for device IQ)
    AcqPlan plan;
    int64_t bytes = 0;
    int slot = -1;
    Clock::time_point t_submit, t_done;
    int state = 0;                  // 0 queued, 1 running, 2 done, 3 failed
    int error = 0;
    std::string message;
    std::vector<ugsdr_acq_result> results;
};

struct Round {
    std::vector<std::shared_ptr<Job>> jobs;
    hipEvent_t start = nullptr, done = nullptr;
    int stream = 0;
    std::pair<int64_t, int64_t> bucket{0, 0};
    int64_t row_blocks = 0;   // sum of P x D x K over the round's requests
};

constexpr int kFirstRoundJobs = 4;

// (transform bucket, surface length): the key requests share a correlation launch by
static std::pair<int64_t, int64_t> bucket_of(const Job& j) { return {j.plan.bucket(), j.plan.L}; }
static int64_t row_blocks(const Job& j) { return j.plan.P * j.plan.D * j.plan.K; }

constexpr size_t kGraphPool = 4;

// an executable graph and the event recorded after its last launch
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
};

// a captured graph handed to the instantiation thread
struct GraphWork {
    std::string key;
    hipGraph_t graph;
};

struct Server {
    ugsdr_hip_ctx* ctx = nullptr;
    ugsdr_server_config cfg{};
    std::vector<hipStream_t> streams;
    std::vector<DevBuf> scratch;  // per stream
    std::vector<ServerSlot> slots;
    std::vector<int> free_slots;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    std::deque<Round> inflight;
    std::vector<hipEvent_t> event_pool;
    uint64_t next_ticket = 1;
    int64_t inflight_bytes = 0;
    int next_stream = 0;
    // the batch-shape autotuner: per bucket, the measured wall time of a
    // round per row-block (exponential average over the bucket's rounds);
    // rounds are cut so that their estimate stays within the budget
    std::map<std::pair<int64_t, int64_t>, double> cost_us;
    double budget_us = 2000.0;
    // round graphs: topology -> executables, shared by the streams (an idle
    // executable's last launch has completed, whichever stream it ran on).
    // The scheduler updates and launches them; the instantiation thread adds
    // new ones, so that hipGraphInstantiate never holds up a round (gmu
    // guards the containers; a deque keeps the scheduler's references valid)
    bool graphs = false;
    // rounds are captured on this stream, never on the streams they run on:
    // the completer queries the running rounds' events meanwhile, which is
    // not permitted on an event whose stream is capturing
    hipStream_t cap_stream = nullptr;
    std::mutex gmu;
    std::condition_variable gcv;
    std::map<std::string, std::deque<GraphExec>> gexec;
    std::map<std::string, int> gpending;   // graphs queued for instantiation
    std::deque<GraphWork> gwork;
    bool gstop = false;
    std::thread ginst;
    std::atomic<int64_t> graph_launches{0}, graph_instantiations{0}, graph_updates{0}, graph_fallbacks{0};
    std::atomic<int64_t> launch_host_ns{0};
    bool stop = false;
    bool sched_done = false;
    ugsdr_server_stats st{};
    std::thread sched, completer;
    int cus = 256;
};

static hipEvent_t take_event(Server& s) {
    if (!s.event_pool.empty()) {
        hipEvent_t e = s.event_pool.back();
        s.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;   // timed: the autotuner reads the rounds' durations
    return e;
}

// A queued or running request pins the replica spectra it reads, so that
// ugsdr_hip_replica_drop refuses to free them under it (caller holds ctx->mu).
static void pin_replicas(ugsdr_hip_ctx* ctx, const std::vector<uint64_t>& keys, int64_t delta) {
    for (uint64_t k : keys) {
        auto it = ctx->replicas.find(k);
        if (it != ctx->replicas.end()) it->second.users += delta;
    }
}

// unpins the replicas of finished jobs (takes ctx->mu; call without s.mu)
static void unpin_jobs(ugsdr_hip_ctx* ctx, const std::vector<std::shared_ptr<Job>>& done) {
    if (done.empty()) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto& j : done) pin_replicas(ctx, j->keys, -1);
}

static void fail_job(Job& j, int rc) {
    j.state = 3;
    j.error = rc;
    j.message = ugsdr_hip_last_error();
    j.t_done = Clock::now();
}

// Topology key of a captured graph: its nodes' types and kernel functions in
// capture order (a stream capture is a chain, so the order fixes the edges).
static bool graph_topology(hipGraph_t g, std::string& key) {
    size_t n = 0;
    if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess) return false;
    std::vector<hipGraphNode_t> nodes(n);
    if (n && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return false;
    key.clear();
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[i], &t) != hipSuccess) return false;
        key.append(reinterpret_cast<const char*>(&t), sizeof t);
        if (t == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            if (hipGraphKernelNodeGetParams(nodes[i], &kp) != hipSuccess) return false;
            key.append(reinterpret_cast<const char*>(&kp.func), sizeof kp.func);
        }
    }
    return true;
}

// Instantiation thread: turns queued captures into executables of their
// topology.
static void instantiate_loop(Server* sp) {
    Server& s = *sp;
    (void)hipSetDevice(s.ctx->device);
    std::unique_lock<std::mutex> lk(s.gmu);
    while (true) {
        s.gcv.wait(lk, [&] { return s.gstop || !s.gwork.empty(); });
        if (s.gwork.empty()) break;   // stopping, nothing left
        GraphWork w = std::move(s.gwork.front());
        s.gwork.pop_front();
        lk.unlock();
        GraphExec e;
        bool ok = hipEventCreateWithFlags(&e.done, hipEventDisableTiming) == hipSuccess;
        if (ok && hipGraphInstantiate(&e.exec, w.graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipEventDestroy(e.done);
            ok = false;
        }
        (void)hipGraphDestroy(w.graph);
        if (!ok) (void)hipGetLastError();
        lk.lock();
        --s.gpending[w.key];
        if (ok) {
            s.gexec[w.key].push_back(e);
            ++s.graph_instantiations;
        }
    }
}

// An idle executable of this topology (its last launch done), updated to the
// captured graph g's parameters; null if there is none.  With none idle and
// room in the pool, g is handed to the instantiation thread (taking it: *g is
// cleared) for a later round of this topology.
static GraphExec* graph_exec_for(Server& s, const std::string& key, hipGraph_t* g) {
    GraphExec* idle = nullptr;
    {
        std::lock_guard<std::mutex> lk(s.gmu);
        auto& pool = s.gexec[key];
        for (auto& e : pool)
            if (e.exec && hipEventQuery(e.done) == hipSuccess) {
                idle = &e;
                break;
            }
        if (!idle) {
            int& pend = s.gpending[key];
            if (pool.size() + (size_t)pend < kGraphPool && !s.gstop) {
                ++pend;
                s.gwork.push_back(GraphWork{key, *g});
                *g = nullptr;
                s.gcv.notify_one();
            }
            return nullptr;
        }
    }
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
    if (hipGraphExecUpdate(idle->exec, *g, &bad, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess) {
        ++s.graph_updates;
        return idle;
    }
    // not updatable after all: retire it (its slot in the pool stays taken)
    (void)hipGetLastError();
    (void)hipGraphExecDestroy(idle->exec);
    idle->exec = nullptr;
    return nullptr;
}

// Runs steps(stream) captured into a graph (on the capture stream) and
// launched on stream si.  False: nothing from steps is on stream si (no
// executable of the round's topology was ready, or the capture failed); the
// caller runs them without a graph.
template <class Steps>
static bool run_as_graph(Server& s, int si, Steps&& steps) {
    hipStream_t st = s.streams[si], cs = s.cap_stream;
    if (!cs || hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    g_capturing = true;
    const int rc = steps(cs);
    g_capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    std::string key;
    GraphExec* e = nullptr;
    if (!rc && ec == hipSuccess && g && graph_topology(g, key)) e = graph_exec_for(s, key, &g);
    if (g) (void)hipGraphDestroy(g);
    if (!e || hipGraphLaunch(e->exec, st) != hipSuccess) {
        (void)hipGetLastError();
        ++s.graph_fallbacks;
        return false;
    }
    (void)hipEventRecord(e->done, st);
    ++s.graph_launches;
    return true;
}

// launches one round (jobs of one bucket) on stream index si; returns rc
static int launch_round(Server& s, Round& rd) {
    hipStream_t st = s.streams[rd.stream];
    UGSDR_HIP_TRY(hipEventRecord(rd.start, st));
    std::vector<const AcqPlan*> plans;
    std::vector<AcqSlot*> devs;
    std::vector<ugsdr_acq_request> reqs(rd.jobs.size());
    // host copies and uploads (outside any graph): each request's IQ into its
    // slot's pinned stage, its tables and IQ to the device, and every slot
    // buffer grown to what the device steps take
    for (size_t i = 0; i < rd.jobs.size(); ++i) {
        Job& j = *rd.jobs[i];
        ServerSlot& sl = s.slots[j.slot];
        ugsdr_acq_request& r = reqs[i];
        r = j.req;
        if (!r.iq_on_device) {
            const size_t bytes = (size_t)j.plan.in_bytes;
            UGSDR_HIP_TRY(sl.iq_stage.reserve(bytes));
            std::memcpy(sl.iq_stage.ptr, j.iq.data(), bytes);
            r.iq = sl.iq_stage.ptr;  // pinned: the async copy does not block this thread
        }
        if (int rc = enqueue_front(j.plan, &r, sl.dev, st, false, FrontStep::Upload)) return rc;
        if (int rc = presize_slot(j.plan, sl.dev, false)) return rc;
        plans.push_back(&j.plan);
        devs.push_back(&sl.dev);
    }
    // the per-stream batch buffers may only grow once the stream's earlier
    // rounds are done; sized for the batch's largest request (the bucket key
    // leaves the block count K free, and the reduced form's scratch grows with K)
    const AcqPlan& p0 = *plans[0];
    size_t need_scratch = 0;
    for (const AcqPlan* p : plans) need_scratch = std::max(need_scratch, corr_scratch_bytes(*p, s.cus));
    if (s.scratch[rd.stream].bytes < need_scratch) {
        UGSDR_HIP_TRY(hipStreamSynchronize(st));
        UGSDR_HIP_TRY(s.scratch[rd.stream].reserve(need_scratch));
    }
    auto device_steps = [&](hipStream_t st) -> int {
        for (size_t i = 0; i < rd.jobs.size(); ++i) {
            const AcqPlan& pl = *plans[i];
            if (int rc = launch_front(pl, &reqs[i], *devs[i], st)) return rc;
            if (int rc = enqueue_classes(pl, *devs[i], st)) return rc;
            if (pl.precision != UGSDR_PREC_FP32)
                if (int rc = enqueue_reduce_classes(pl, *devs[i], st)) return rc;
        }
        if (int rc = p0.precision == UGSDR_PREC_FP32
                         ? enqueue_corr(plans, devs, st, s.scratch[rd.stream], s.cus, p0.threads)
                         : enqueue_corr_reduced(plans, devs, st, s.scratch[rd.stream], s.cus, p0.threads))
            return rc;
        for (size_t i = 0; i < rd.jobs.size(); ++i)
            if (int rc = enqueue_final(*plans[i], *devs[i], st, false)) return rc;
        return UGSDR_OK;
    };
    if (!(s.graphs && run_as_graph(s, rd.stream, device_steps)))
        if (int rc = device_steps(st)) return rc;
    for (auto& jp : rd.jobs) {
        Job& j = *jp;
        ServerSlot& sl = s.slots[j.slot];
        const size_t rb = (size_t)j.plan.P * sizeof(ugsdr_acq_result);
        UGSDR_HIP_TRY(sl.results.reserve(rb));
        UGSDR_HIP_TRY(hipMemcpyAsync(sl.results.ptr, sl.dev.results.ptr, rb, hipMemcpyDeviceToHost, st));
    }
    UGSDR_HIP_TRY(hipEventRecord(rd.done, st));
    return UGSDR_OK;
}

static void scheduler_loop(Server* sp) {
    Server& s = *sp;
    (void)hipSetDevice(s.ctx->device);
    std::unique_lock<std::mutex> lk(s.mu);
    while (true) {
        s.cv_sched.wait(lk, [&] { return s.stop || (!s.queue.empty() && !s.free_slots.empty()); });
        if (s.stop && s.queue.empty()) {
            s.sched_done = true;
            break;
        }
        // admission in arrival order
        std::vector<std::shared_ptr<Job>> admitted;
        while (!s.queue.empty() && !s.free_slots.empty()) {
            auto& j = s.queue.front();
            if (s.inflight_bytes + j->bytes > s.cfg.max_device_bytes && s.inflight_bytes > 0) break;
            j->slot = s.free_slots.back();
            s.free_slots.pop_back();
            s.inflight_bytes += j->bytes;
            j->state = 1;
            admitted.push_back(j);
            s.queue.pop_front();
        }
        if (admitted.empty()) {
            // budget exhausted: wait for a completion to free memory
            s.cv_sched.wait_for(lk, std::chrono::microseconds(200));
            continue;
        }
        // group by bucket; cut each group into rounds of at most max_batch
        // requests whose estimated time (row-blocks x the bucket's measured
        // cost) stays within the budget.  A bucket without a measurement yet
        // runs rounds of at most kFirstRoundJobs requests until its first
        // round completes.
        std::map<std::pair<int64_t, int64_t>, std::vector<std::shared_ptr<Job>>> groups;
        for (auto& j : admitted) groups[bucket_of(*j)].push_back(j);
        std::vector<Round> rounds;
        for (auto& g : groups) {
            const auto it = s.cost_us.find(g.first);
            const double cost = it == s.cost_us.end() ? -1.0 : it->second;
            Round rd;
            double est = 0.0;
            auto close = [&]() {
                rd.bucket = g.first;
                rd.start = take_event(s);
                rd.done = take_event(s);
                rd.stream = s.next_stream;
                s.next_stream = (s.next_stream + 1) % (int)s.streams.size();
                rounds.push_back(std::move(rd));
                rd = Round();
                est = 0.0;
            };
            for (auto& j : g.second) {
                const double e = cost < 0 ? 0.0 : cost * (double)row_blocks(*j);
                const int cap = cost < 0 ? std::min(s.cfg.max_batch, kFirstRoundJobs) : s.cfg.max_batch;
                if (!rd.jobs.empty() && ((int)rd.jobs.size() >= cap || est + e > s.budget_us))
                    close();
                rd.jobs.push_back(j);
                rd.row_blocks += row_blocks(*j);
                est += e;
            }
            if (!rd.jobs.empty()) close();
        }
        lk.unlock();
        std::vector<std::shared_ptr<Job>> failed;
        for (auto& rd : rounds) {
            const auto t_launch = Clock::now();
            int rc = rd.done && rd.start ? launch_round(s, rd) : UGSDR_ERR_HIP;
            s.launch_host_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t_launch).count();
            std::lock_guard<std::mutex> g(s.mu);
            if (rc) {
                // the round's work may be partly queued: wait for the stream, then fail its jobs
                (void)hipStreamSynchronize(s.streams[rd.stream]);
                for (auto& j : rd.jobs) {
                    fail_job(*j, rc);
                    s.free_slots.push_back(j->slot);
                    s.inflight_bytes -= j->bytes;
                    ++s.st.failed;
                    failed.push_back(j);
                }
                if (rd.done) s.event_pool.push_back(rd.done);
                if (rd.start) s.event_pool.push_back(rd.start);
                s.cv_done.notify_all();
                continue;
            }
            ++s.st.rounds;
            ++s.st.batches;
            s.st.max_batch_seen = std::max<int64_t>(s.st.max_batch_seen, (int64_t)rd.jobs.size());
            s.inflight.push_back(std::move(rd));
        }
        unpin_jobs(s.ctx, failed);
        lk.lock();
    }
}

static void completer_loop(Server* sp) {
    Server& s = *sp;
    (void)hipSetDevice(s.ctx->device);
    while (true) {
        bool any = false;
        std::vector<std::shared_ptr<Job>> finished;
        {
            std::unique_lock<std::mutex> lk(s.mu);
            if (s.sched_done && s.inflight.empty()) break;
            for (auto it = s.inflight.begin(); it != s.inflight.end();) {
                const hipError_t q = hipEventQuery(it->done);
                if (q == hipErrorNotReady) {
                    ++it;
                    continue;
                }
                const auto now = Clock::now();
                for (auto& jp : it->jobs) {
                    Job& j = *jp;
                    if (q == hipSuccess) {
                        j.results.resize(j.plan.P);
                        std::memcpy(j.results.data(), s.slots[j.slot].results.ptr,
                                    (size_t)j.plan.P * sizeof(ugsdr_acq_result));
                        j.state = 2;
                        j.t_done = now;
                        ++s.st.completed;
                    } else {
                        set_error(std::string("round failed: ") + hipGetErrorString(q));
                        fail_job(j, UGSDR_ERR_HIP);
                        ++s.st.failed;
                    }
                    s.free_slots.push_back(j.slot);
                    s.inflight_bytes -= j.bytes;
                    finished.push_back(jp);
                }
                if (q == hipSuccess && it->row_blocks > 0) {
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, it->start, it->done) == hipSuccess && ms > 0.f) {
                        const double c = 1000.0 * ms / (double)it->row_blocks;
                        auto& e = s.cost_us[it->bucket];
                        e = e > 0.0 ? 0.75 * e + 0.25 * c : c;
                    }
                }
                s.event_pool.push_back(it->done);
                s.event_pool.push_back(it->start);
                it = s.inflight.erase(it);
                any = true;
            }
            if (any) {
                s.cv_done.notify_all();
                s.cv_sched.notify_all();
            }
        }
        unpin_jobs(s.ctx, finished);
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace ugsdr_hip

using namespace ugsdr_hip;

struct ugsdr_hip_server {
    Server s;
};

extern "C" {

int ugsdr_hip_server_create(ugsdr_hip_ctx* ctx, const ugsdr_server_config* cfg, ugsdr_hip_server** out) {
    if (!ctx || !out) {
        set_error("null argument");
        return UGSDR_ERR_ARG;
    }
    *out = nullptr;
    UGSDR_HIP_TRY(hipSetDevice(ctx->device));
    auto* h = new ugsdr_hip_server();
    Server& s = h->s;
    s.ctx = ctx;
    if (cfg) s.cfg = *cfg;
    if (s.cfg.max_device_bytes <= 0) s.cfg.max_device_bytes = 8ll << 30;
    if (s.cfg.max_inflight <= 0) s.cfg.max_inflight = 64;
    if (s.cfg.n_streams <= 0) s.cfg.n_streams = 4;
    if (s.cfg.max_batch <= 0 || s.cfg.max_batch > kMaxBatch) s.cfg.max_batch = kMaxBatch;
    if (s.cfg.round_budget_us > 0) s.budget_us = (double)s.cfg.round_budget_us;
    s.graphs = s.cfg.round_graphs != 0;
    s.cus = device_cu_count(ctx->device);
    s.streams.resize(s.cfg.n_streams);
    for (auto& st : s.streams) {
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);  // default priority
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            for (auto& x : s.streams)
                if (x) (void)hipStreamDestroy(x);
            delete h;
            return UGSDR_ERR_HIP;
        }
    }
    // without a capture stream the rounds run eagerly
    if (s.graphs && hipStreamCreateWithFlags(&s.cap_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s.cap_stream = nullptr;
        s.graphs = false;
    }
    s.scratch.resize(s.cfg.n_streams);
    if (s.graphs) s.ginst = std::thread(instantiate_loop, &s);
    s.slots.resize(s.cfg.max_inflight);
    for (int i = s.cfg.max_inflight - 1; i >= 0; --i) s.free_slots.push_back(i);
    s.sched = std::thread(scheduler_loop, &s);
    s.completer = std::thread(completer_loop, &s);
    *out = h;
    return UGSDR_OK;
}

int ugsdr_hip_server_submit(ugsdr_hip_server* h, const ugsdr_acq_request* req, uint64_t* ticket) {
    if (!h || !req || !ticket) {
        set_error("null argument");
        return UGSDR_ERR_ARG;
    }
    Server& s = h->s;
    auto j = std::make_shared<Job>();
    j->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s.ctx->mu);
        if (int rc = make_acq_plan(s.ctx, req, j->plan)) return rc;
        if (req->precision != UGSDR_PREC_FP32 && !reduced_plan_ok(j->plan)) {
            set_error("no reduced-precision plan for n_fft " + std::to_string(j->plan.N) + ", surface_len " +
                      std::to_string(j->plan.L));
            return UGSDR_ERR_UNSUPPORTED;
        }
        // the FP32 upsampled-code search has nothing to tune; the reduced
        // form takes a pinned variant or its default, as ugsdr_hip_acquire
        const auto pin = s.ctx->tuned.find(j->plan.N);
        j->plan.threads = req->precision != UGSDR_PREC_FP32 ? (pin != s.ctx->tuned.end() ? pin->second : 1024)
                          : j->plan.ucode                   ? 0
                                                            : tuned_threads(s.ctx, j->plan.N);
    }
    j->req = *req;
    j->doppler.assign(req->doppler, req->doppler + req->n_doppler);
    j->keys.assign(req->replica_keys, req->replica_keys + req->n_prn);
    j->req.doppler = j->doppler.data();
    j->req.replica_keys = j->keys.data();
    if (!req->iq_on_device) {
        const size_t bytes = (size_t)j->plan.in_bytes;
        if (j->plan.fmt == UGSDR_FMT_BBP_DDC)
            if (int rc = verify_bbp_headers(req->iq, j->plan.in_bytes / j->plan.epoch_bytes, j->plan.epoch_bytes))
                return rc;
        j->iq.assign(static_cast<const char*>(req->iq), static_cast<const char*>(req->iq) + bytes);
        j->req.iq = j->iq.data();
    }
    j->bytes = j->plan.footprint(false);
    {
        // the keys were resolved above; a key dropped since then fails here
        std::lock_guard<std::mutex> lk(s.ctx->mu);
        for (uint64_t k : j->keys)
            if (!s.ctx->replicas.count(k)) {
                set_error("unknown replica key");
                return UGSDR_ERR_STATE;
            }
        pin_replicas(s.ctx, j->keys, +1);
    }
    {
        std::lock_guard<std::mutex> lk(s.mu);
        int rc = UGSDR_OK;
        if (s.stop) {
            set_error("server is shutting down");
            rc = UGSDR_ERR_STATE;
        } else if (j->bytes > s.cfg.max_device_bytes) {
            set_error("request exceeds the server's device-memory budget");
            rc = UGSDR_ERR_NOMEM;
        }
        if (rc) {
            std::lock_guard<std::mutex> lc(s.ctx->mu);
            pin_replicas(s.ctx, j->keys, -1);
            return rc;
        }
        j->ticket = s.next_ticket++;
        s.jobs[j->ticket] = j;
        s.queue.push_back(j);
        ++s.st.submitted;
    }
    s.cv_sched.notify_one();
    *ticket = j->ticket;
    return UGSDR_OK;
}

int ugsdr_hip_server_poll(ugsdr_hip_server* h, uint64_t ticket, int wait_ms, ugsdr_acq_result* out,
                          double* latency_ms) {
    if (!h) {
        set_error("null server");
        return UGSDR_ERR_ARG;
    }
    Server& s = h->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.jobs.find(ticket);
    if (it == s.jobs.end()) {
        set_error("unknown ticket");
        return UGSDR_ERR_STATE;
    }
    auto j = it->second;
    if (j->state < 2 && wait_ms != 0) {
        auto pred = [&] { return j->state >= 2; };
        if (wait_ms < 0) s.cv_done.wait(lk, pred);
        else s.cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    if (j->state < 2) return UGSDR_PENDING;
    s.jobs.erase(ticket);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    if (j->state == 3) {
        set_error(j->message);
        return j->error;
    }
    if (out) {
        if (j->req.hypotheses > 1)
            reduce_hypotheses(j->results.data(), (int64_t)j->results.size(), j->req.hypotheses, j->req.threshold, out);
        else
            std::memcpy(out, j->results.data(), j->results.size() * sizeof(ugsdr_acq_result));
    }
    return UGSDR_OK;
}

int ugsdr_hip_server_stats(ugsdr_hip_server* h, ugsdr_server_stats* out) {
    if (!h || !out) {
        set_error("null argument");
        return UGSDR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(h->s.mu);
    *out = h->s.st;
    out->inflight = (int64_t)(h->s.cfg.max_inflight - (int)h->s.free_slots.size());
    out->inflight_bytes = h->s.inflight_bytes;
    out->graph_launches = h->s.graph_launches;
    out->graph_instantiations = h->s.graph_instantiations;
    out->graph_updates = h->s.graph_updates;
    out->graph_fallbacks = h->s.graph_fallbacks;
    out->launch_host_us = h->s.launch_host_ns / 1000;
    return UGSDR_OK;
}

int ugsdr_hip_server_destroy(ugsdr_hip_server* h) {
    if (!h) return UGSDR_OK;
    Server& s = h->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
    }
    s.cv_sched.notify_all();
    if (s.sched.joinable()) s.sched.join();
    if (s.completer.joinable()) s.completer.join();
    (void)hipSetDevice(s.ctx->device);
    for (auto& st : s.streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    {
        std::lock_guard<std::mutex> lk(s.gmu);
        s.gstop = true;
    }
    s.gcv.notify_all();
    if (s.ginst.joinable()) s.ginst.join();   // finishes the queued instantiations first
    if (s.cap_stream) (void)hipStreamDestroy(s.cap_stream);
    for (auto e : s.event_pool) (void)hipEventDestroy(e);
    for (auto& kv : s.gexec)
        for (auto& e : kv.second) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.done) (void)hipEventDestroy(e.done);
        }
    for (auto& sl : s.slots) sl.dev.release_all();
    for (auto& d : s.scratch) d.release();
    delete h;
    return UGSDR_OK;
}

}  // extern "C"
