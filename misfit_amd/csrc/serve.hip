// serve.hip -- the serving layer (host code only): many graphs held on one GPU, update
// batches taken as they arrive.
//
// The reference's misfit_async overlaps the host work of batch i+1 with the
// device work of batch i for one graph (src/misfit_async/misfit_async.cu:
// 484-701,827-876).  This generalises it to continuous batching:
//   * graphs are admitted by free HBM (their device footprint must fit in
//     the free memory minus a reserve), each in its own context and stream;
//   * requests (update batches) are queued per graph in arrival order; a
//     pool of dispatcher workers, each owning a share of the graphs, keeps
//     up to max_inflight batches of every graph queued on the device, so
//     batches of different graphs overlap on the GPU and a request that
//     arrives while others run starts as soon as its graph has a free slot;
//     a batch whose buffers may not fit in free HBM is refused
//     (MISFIT_E_NOMEM) instead of being started;
//   * with max_inflight 1 every batch is queued with the rounds it is
//     expected to need and gets any further rounds from its worker when its
//     tail halts, so no graph's rounds run empty on the device;
//   * a graph whose overlay passes compact_fraction of its base entries is
//     compacted once its in-flight batches have drained (its later requests
//     wait); every request reports how many compactions of its graph ran
//     before it started, which places each compaction exactly;
//   * every request reports its queueing time, its latency from submission
//     to result and its device time.
// Kernel launches cost the host about 4 us each and scale with threads on
// separate streams, so several workers feed the GPU; each graph's context
// is only ever called by its worker (or, while no worker is inside it, by a
// public call holding the server lock).  The workers take the server lock
// for queue changes only, never around a HIP call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/misfit.h"
#include "graph.h"

// capi.hip: 1 while a batch of the context may still need rounds queued
// from the host (queued without a reliable round estimate).
int mf_ctx_settling(const misfit_ctx* c);
// capi.hip: checks and prepares a batch for the serving kernel: with
// `whole` nothing is launched, else its prologue is queued on the graph's
// stream and mf_ctx_external_ready reports when it has ended.
int mf_ctx_prepare_external(misfit_ctx* c, const misfit_update* ups, int64_t count, bool whole, uint64_t* ticket,
                            BP* P);
int mf_ctx_external_ready(misfit_ctx* c, uint64_t ticket);
// capi.hip: every batch of the context may halt (one batch per graph).
void mf_ctx_set_halting(misfit_ctx* c, bool on);
// capi.hip: whether compactions keep their released buffers for the next
// one, and how many bytes a context keeps.
void mf_ctx_set_keep_compact(misfit_ctx* c, bool on);
int64_t mf_ctx_spare_bytes(const misfit_ctx* c);

namespace {
using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

enum ReqState { PENDING = 0, INFLIGHT = 1, DONE = 2 };

struct Request {
    uint64_t id = 0;
    int32_t graph = -1;
    std::vector<misfit_update> ups;
    clk::time_point t_submit, t_enqueue;
    uint64_t ticket = 0;
    bool external = false;  // runs in the serving kernel, not on the graph's stream
    bool published = false; // ... and its item is in the kernel's queue (after its prologue)
    BP P{};                 // ... its parameters, until published
    ReqState state = PENDING;
    int rc = MISFIT_OK;
    std::string err;
    misfit_request_stats st{};
};

struct Graph {
    int32_t id = -1;
    int worker = 0;
    misfit_ctx* ctx = nullptr;
    int32_t n = 0;
    int64_t base_m = 0;
    std::deque<Request*> pending, inflight;
    bool compact_pending = false;
    bool in_call = false;   // a worker is calling the context outside the lock
    int64_t compactions = 0;
    int64_t overlay = 0;            // overlay pool slots after the last finished batch
    int64_t compact_at = 0;         // overlay size that triggers the next compaction (0: by the fraction)
    int compact_rc = MISFIT_OK;     // a failed compaction's error, for the next request to start
    std::string compact_err;
    double last_epr = 0.0;          // edges the graph's last batch committed per round (0: no batch yet)
    int64_t last_window = 0;        // ... the edges it speculated per round
    bool last_kernel = false;       // ... and whether it ran in the serving kernel
    int external = 0;               // its batches running in the serving kernel
};

// The serving kernel's host side: the pinned queue (coherent, so the
// kernel sees each item and the raised tail without a copy), the device
// head counter the workgroups claim items with, the kernel's stream and
// the end event of its latest launch.  A launch ends on its own once no
// item comes for the idle time; a caller whose item has not finished calls
// keep_running, which launches again when the last launch has ended (a
// claimed item always finishes before its workgroup leaves, so an
// unfinished item after the end was never claimed).
struct Runner {
    ServeQueue* q = nullptr;
    ServeQueue* q_dev = nullptr;
    u64* d_head = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ended = nullptr;
    bool launched = false;
    u64 pushed = 0, finished = 0;
    int cus = 0;
    int threads = 1024;  // per workgroup (MISFIT_SERVE_THREADS: 256, 512 or 1024)
    u64 idle_ticks = 0;
    int err = MISFIT_OK;
    std::string err_msg;
    std::mutex mu;

    int init(int device, double idle_ms) {
        hipSetDevice(device);
        int khz = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
            khz = 100000;
        idle_ticks = (u64)(idle_ms * (double)khz);
        const char* th = getenv("MISFIT_SERVE_THREADS");
        if (th) threads = atoi(th) >= 1024 ? 1024 : atoi(th) >= 512 ? 512 : 256;
        void* dp = nullptr;
        if (hipHostMalloc((void**)&q, sizeof(ServeQueue), hipHostMallocCoherent) != hipSuccess) return MISFIT_E_NOMEM;
        memset((void*)q, 0, sizeof(ServeQueue));
        if (hipHostGetDevicePointer(&dp, q, 0) != hipSuccess) return MISFIT_E_HIP;
        q_dev = (ServeQueue*)dp;
        if (hipMalloc((void**)&d_head, sizeof(u64)) != hipSuccess) return MISFIT_E_NOMEM;
        // A stream with a CU mask gets a hardware queue of its own: the
        // long-running kernel must not hold up the graphs' streams, which
        // share the other hardware queues (a queue runs its kernels in
        // order).  The mask covers every CU.
        std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
        for (int cu = 0; cu < cus; cu++) mask[(size_t)cu / 32] |= 1u << (cu % 32);
        if (hipExtStreamCreateWithCUMask(&stream, (uint32_t)mask.size(), mask.data()) != hipSuccess)
            return MISFIT_E_HIP;
        // on the kernel's own stream: ordered before its first launch
        if (hipMemsetAsync(d_head, 0, sizeof(u64), stream) != hipSuccess) return MISFIT_E_HIP;
        if (hipStreamSynchronize(stream) != hipSuccess) return MISFIT_E_HIP;
        if (hipEventCreateWithFlags(&ended, hipEventDisableTiming) != hipSuccess) return MISFIT_E_HIP;
        return MISFIT_OK;
    }
    // Relaunches the kernel (grid: one workgroup per graph that may have a
    // batch in it, at most one per CU) when its last launch has ended.
    int keep_running(int device, int graphs) {
        std::lock_guard<std::mutex> lk(mu);
        if (err) return err;
        if (launched) {
            const hipError_t e = hipEventQuery(ended);
            if (e == hipErrorNotReady) return MISFIT_OK;
            if (e != hipSuccess) return fail_locked(MISFIT_E_HIP, std::string("serving kernel: ") + hipGetErrorString(e));
            launched = false;
            mf_serving_end();
        }
        if (finished >= pushed) return MISFIT_OK;  // nothing left for a kernel to do
        hipSetDevice(device);
        // workgroups of 512 or 1024 threads take a CU each, of 256 half of one
        const int grid = std::max(1, std::min(threads <= 256 ? 2 * cus : cus, graphs));
        mf_serving_begin();  // frees wait from now on until the launch has ended
        mf_launch_serve(q_dev, d_head, grid, threads, idle_ticks, stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(ended, stream);
        if (e != hipSuccess) {
            mf_serving_end();
            return fail_locked(MISFIT_E_HIP, std::string("serving kernel launch: ") + hipGetErrorString(e));
        }
        launched = true;
        return MISFIT_OK;
    }
    // Publishes one batch; the queue holds at most MF_SERVE_CAP unfinished.
    int push(const BP& P, bool whole, int device, int graphs) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (err) return err;
            if (pushed - finished >= (u64)MF_SERVE_CAP) return MISFIT_E_STATE;
            memcpy((void*)&q->items[pushed % MF_SERVE_CAP].P, &P, sizeof(BP));
            q->items[pushed % MF_SERVE_CAP].whole = whole ? 1 : 0;
            pushed++;
            __atomic_store_n(&q->tail, pushed, __ATOMIC_RELEASE);
        }
        return keep_running(device, graphs);
    }
    std::string error() {
        std::lock_guard<std::mutex> lk(mu);
        return err_msg;
    }
    void done_one() {
        std::lock_guard<std::mutex> lk(mu);
        finished++;
    }
    int fail_locked(int rc, const std::string& m) {
        err = rc;
        err_msg = m;
        return rc;
    }
    void shutdown() {
        if (!q) return;
        __atomic_store_n(&q->stop, 1ull, __ATOMIC_RELEASE);
        if (launched) {
            hipEventSynchronize(ended);
            mf_serving_end();
        }
        launched = false;
        if (ended) hipEventDestroy(ended);
        if (stream) hipStreamDestroy(stream);
        if (d_head) hipFree(d_head);
        hipHostFree(q);
        q = nullptr;
    }
};

// Device bytes a resident graph may hold: the base (int64 offsets, int32
// entries), the sorted copy of its rows (present after any compaction, or
// for unsorted input), about m more for the overlay pool, the inserted-pair
// hash and the row-chunk tasks, the per-vertex state and compaction scratch
// (about 256 bytes a vertex), and the first batch buffers.
int64_t graph_footprint(int32_t n, int64_t m) {
    return 8 * ((int64_t)n + 1) + 4 * m + 8 * m + m + 256 * (int64_t)n + (int64_t(64) << 20);
}

// Device bytes a compaction of that graph allocates while it runs: the new
// offsets and entries and their sorted copy (12 bytes an entry, for a base
// grown by the compaction fraction plus a tenth), the old rows' deletion
// marks in position and in sorted order (16 bytes per 64 entries each), the
// per-row lists and the library sort's scratch.
int64_t compaction_headroom(int32_t n, int64_t m, double fraction) {
    const double grown = (double)m * (1.1 + (fraction > 0 ? fraction : 0.0));
    return 12 * (int64_t)grown + m / 2 + 24 * ((int64_t)n + 1) + (int64_t(16) << 20);
}

// Device bytes one batch of `count` updates may add (batch buffers doubled
// to the new capacity, overlay pool and inserted-pair hash growth).
int64_t batch_footprint(int64_t count) { return 1024 * count; }

int64_t device_free_bytes(int device) {
    size_t fr = 0, tot = 0;
    hipSetDevice(device);
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return (int64_t)fr;
}
}  // namespace

struct misfit_server {
    int device = 0;
    misfit_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<int32_t, std::unique_ptr<Graph>> graphs;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_req = 1;
    int32_t next_graph = 0;
    int64_t loading_bytes = 0;  // footprints of graphs being loaded (admitted, not yet listed)
    bool stop = false;
    std::vector<std::thread> workers;
    std::string err;
    bool compacting = false;    // one compaction at a time: each needs its headroom only while it runs
    int64_t headroom_max = 0;   // the largest compaction headroom among the resident graphs
    Runner runner;              // the serving kernel (opts.persistent)
    bool check_each = false;    // MISFIT_SERVE_CHECK=1: the full check after every batch (a debugging aid)
    // MISFIT_SERVE_COMPACT_KEEP=1: every graph keeps the buffers its
    // compactions release for its next one (no free and allocation of the
    // base and its sorted copy per compaction); admission then charges each
    // graph its own compaction headroom instead of sharing the largest one
    bool keep_compact = false;

    // Whether request r of graph g runs its rounds in the serving kernel:
    // every batch (persistent 2), or (persistent 1) a small one, or one of a
    // contended graph, whose last batch committed few edges per round: fewer
    // than 256 on the graph's stream (its grid-wide rounds leave most of the
    // GPU idle and cost three launches each), or under three quarters of
    // the window in the kernel (a wider grid-wide window would not commit
    // more).  The rest, and a graph's first large batch, run on its stream.
    bool use_kernel(const Graph* g, const Request* r) const {
        if (opts.persistent <= 0 || opts.max_inflight != 1) return false;
        if (opts.persistent >= 2 || r->ups.size() <= 4096) return true;
        if (g->last_epr <= 0.0) return false;
        if (g->last_kernel) return g->last_epr < 0.75 * (double)g->last_window;
        return g->last_epr < 256.0;
    }
    // ... and whether the kernel runs its prologue too: tiny batches, whose
    // prologue kernels would each cost more to launch than to run.
    static bool whole_in_kernel(const Request* r) { return r->ups.size() <= 1024; }

    int fail(int rc, const std::string& m) {
        err = m;
        return rc;
    }
    bool busy(int w) const {
        for (auto& kv : graphs)
            if (kv.second->worker == w && (!kv.second->pending.empty() || !kv.second->inflight.empty())) return true;
        return false;
    }
    bool inflight(int w) const {
        for (auto& kv : graphs)
            if (kv.second->worker == w && !kv.second->inflight.empty()) return true;
        return false;
    }
    void finish(Request* r, int rc, const std::string& msg) {
        r->rc = rc;
        r->err = msg;
        r->state = DONE;
        r->st.latency_ms = ms_between(r->t_submit, clk::now());
        cv_done.notify_all();
    }
    Graph* graph(int32_t id) {
        auto it = graphs.find(id);
        return it == graphs.end() ? nullptr : it->second.get();
    }
    bool service(Graph* g, std::unique_lock<std::mutex>& lk);
    void run(int w);
};

// One pass over graph g by its worker: collect finished batches, compact
// if due, start queued requests.  Called and returns with the lock held;
// every context call runs without it.  Returns whether anything happened.
bool misfit_server::service(Graph* g, std::unique_lock<std::mutex>& lk) {
    bool progress = false;
    g->in_call = true;
    while (!g->inflight.empty()) {
        Request* r = g->inflight.front();
        const int ngraphs = (int)graphs.size();
        lk.unlock();
        int p = 0;
        std::string em;
        if (r->external && !r->published) {  // its prologue on the stream first, then the kernel
            p = mf_ctx_external_ready(g->ctx, r->ticket);
            if (p < 0) {
                em = misfit_last_error(g->ctx);
            } else if (p == 1) {
                // a batch whose prologue ran but whose rounds cannot (the
                // kernel failed to launch) leaves its graph unrepaired
                p = runner.push(r->P, false, device, ngraphs);
                if (p == MISFIT_OK) r->published = true;
                else em = "serving kernel unavailable: " + runner.error();
            }
        } else {
            p = misfit_poll(g->ctx, r->ticket);
            if (p < 0) em = misfit_last_error(g->ctx);
        }
        misfit_batch_stats bs{};
        int rc = p < 0 ? p : MISFIT_OK;
        if (p == 0 && r->external && r->published) {  // not finished: the kernel must still be running
            rc = runner.keep_running(device, ngraphs);
            if (rc != MISFIT_OK) {
                p = rc;
                em = "serving kernel unavailable: " + runner.error();
            }
        }
        if (p == 1) {
            rc = misfit_wait(g->ctx, r->ticket, &bs);
            if (rc != MISFIT_OK) em = misfit_last_error(g->ctx);
            if (r->external) runner.done_one();
            if (rc == MISFIT_OK && check_each) {  // MISFIT_SERVE_CHECK=1: the full check after every batch
                int64_t cf = 0, un = 0, ce = 0;
                rc = misfit_check(g->ctx, &cf, &un, &ce);
                if (rc == MISFIT_OK && (cf || un || ce)) {
                    rc = MISFIT_E_CHECK;
                    em = "check after the batch failed (" + std::string(r->external ? "serving kernel" : "stream") +
                         ", " + std::to_string(r->ups.size()) + " updates, " + std::to_string(g->compactions) +
                         " compactions so far): conflicts " + std::to_string(cf) + ", uncovered " +
                         std::to_string(un) + ", count errors " + std::to_string(ce);
                }
            }
        }
        lk.lock();
        if (p == 0) break;
        g->inflight.pop_front();
        if (r->external) g->external--;
        progress = true;
        if (rc != MISFIT_OK) {
            finish(r, rc, em);
            continue;
        }
        if (bs.rounds > 0) {
            g->last_epr = (double)bs.updates / (double)bs.rounds;
            g->last_window = bs.window;
            g->last_kernel = r->external;
        }
        r->st.batch = bs;
        r->st.device_ms = bs.device_ms;
        g->overlay = bs.overlay_entries;
        const double due = g->compact_at > 0 ? (double)g->compact_at : opts.compact_fraction * (double)g->base_m;
        if (opts.compact_fraction > 0 && (double)bs.overlay_entries > due) g->compact_pending = true;
        finish(r, MISFIT_OK, "");
    }
    // compactions run one at a time (each needs its headroom only while it
    // runs); a graph whose turn has not come keeps its requests waiting
    if (g->compact_pending && g->inflight.empty() && !compacting) {
        compacting = true;
        lk.unlock();
        int64_t new_m = 0;
        int rc = MISFIT_OK;
        std::string em;
        // a graph that keeps its buffers needs only what they do not cover
        const int64_t need =
            std::max<int64_t>(0, compaction_headroom(g->n, g->base_m + g->overlay, opts.compact_fraction) -
                                     (keep_compact ? mf_ctx_spare_bytes(g->ctx) : 0));
        const int64_t fr = device_free_bytes(device);
        if (fr - need < opts.hbm_reserve) {
            rc = MISFIT_E_NOMEM;
            em = "compaction not started: needs " + std::to_string(need >> 20) + " MiB, " + std::to_string(fr >> 20) +
                 " MiB free, " + std::to_string(opts.hbm_reserve >> 20) + " MiB reserved";
        } else {
            rc = misfit_compact(g->ctx, &new_m);
            if (rc != MISFIT_OK) em = std::string("compaction failed: ") + misfit_last_error(g->ctx);
        }
        lk.lock();
        compacting = false;
        g->compact_pending = false;
        if (rc == MISFIT_OK) {
            g->base_m = new_m;
            g->overlay = 0;
            g->compact_at = 0;
            g->compactions++;
        } else {
            // the graph is unchanged; the next request to start reports the
            // error, and the compaction is tried again only once the overlay
            // has grown by another compaction fraction of the base
            g->compact_rc = rc;
            g->compact_err = em;
            g->compact_at = g->overlay + (int64_t)(opts.compact_fraction * (double)g->base_m) + 1;
        }
        progress = true;
        cv_work.notify_all();  // another graph may be waiting for its turn to compact
    }
    // a graph whose batch is still settling its round count gets no
    // further batch until it finishes (the poll above queues its rounds)
    while (!g->compact_pending && (int)g->inflight.size() < opts.max_inflight && !g->pending.empty() &&
           !mf_ctx_settling(g->ctx)) {
        Request* r = g->pending.front();
        g->pending.pop_front();
        progress = true;
        if (g->compact_rc != MISFIT_OK) {  // a failed compaction's error goes to the next request
            const int rc = g->compact_rc;
            g->compact_rc = MISFIT_OK;
            finish(r, rc, g->compact_err);
            continue;
        }
        const bool ext = use_kernel(g, r);
        const int ngraphs = (int)graphs.size();
        lk.unlock();
        int rc = MISFIT_OK;
        std::string em;
        if (device_free_bytes(device) - batch_footprint((int64_t)r->ups.size()) < opts.hbm_reserve) {
            rc = MISFIT_E_NOMEM;
            em = "batch not admitted: not enough free device memory";
        } else if (ext) {
            const bool whole = whole_in_kernel(r);
            rc = mf_ctx_prepare_external(g->ctx, r->ups.data(), (int64_t)r->ups.size(), whole, &r->ticket, &r->P);
            if (rc != MISFIT_OK) {
                em = misfit_last_error(g->ctx);
            } else if (whole) {
                rc = runner.push(r->P, true, device, ngraphs);
                if (rc == MISFIT_OK) r->published = true;
                else em = "serving kernel unavailable: " + runner.error();
            }
        } else {
            rc = misfit_submit_batch(g->ctx, r->ups.data(), (int64_t)r->ups.size(), &r->ticket);
            if (rc != MISFIT_OK) em = misfit_last_error(g->ctx);
        }
        lk.lock();
        if (rc != MISFIT_OK) {
            finish(r, rc, em);
            continue;
        }
        r->external = ext;
        if (ext) g->external++;
        r->t_enqueue = clk::now();
        r->st.compactions_before = g->compactions;
        r->st.queue_ms = ms_between(r->t_submit, r->t_enqueue);
        r->state = INFLIGHT;
        g->inflight.push_back(r);
    }
    g->in_call = false;
    cv_done.notify_all();  // public calls waiting for the context
    return progress;
}

void misfit_server::run(int w) {
    std::unique_lock<std::mutex> lk(mu);
    std::vector<int32_t> mine;
    while (true) {
        if (stop && !busy(w)) break;
        mine.clear();
        for (auto& kv : graphs)
            if (kv.second->worker == w) mine.push_back(kv.first);
        bool progress = false;
        for (int32_t id : mine) {
            Graph* g = graph(id);  // looked up again: the lock was released in between
            if (g) progress |= service(g, lk);
        }
        if (progress) continue;
        if (!inflight(w)) {
            cv_work.wait(lk);  // until a submit, a new graph or the stop request
        } else {
            // batches on the device: look again shortly (submits wake us too)
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

extern "C" int misfit_server_create(int device, const misfit_server_opts* opts, misfit_server** out) {
    if (!out) return MISFIT_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MISFIT_E_ARG;
    misfit_server* s = new misfit_server();
    s->device = device;
    s->opts.hbm_reserve = opts ? opts->hbm_reserve : 0;
    s->opts.compact_fraction = opts ? opts->compact_fraction : 0.05;
    int mi = opts ? opts->max_inflight : 2;
    s->opts.max_inflight = mi < 1 ? 1 : (mi > 3 ? 3 : mi);  // a context stages at most 3 batches ahead
    int nw = opts ? opts->workers : 0;
    if (nw <= 0) nw = 8;
    s->opts.workers = nw > 64 ? 64 : nw;
    s->opts.persistent = opts ? opts->persistent : 0;
    const char* ck = getenv("MISFIT_SERVE_CHECK");
    s->check_each = ck && ck[0] == '1';
    const char* kc = getenv("MISFIT_SERVE_COMPACT_KEEP");
    s->keep_compact = kc && kc[0] == '1';
    if (s->opts.persistent > 0) {
        const char* idle = getenv("MISFIT_SERVE_IDLE_MS");
        const int rc = s->runner.init(device, idle && atof(idle) > 0 ? atof(idle) : 20.0);
        if (rc != MISFIT_OK) {
            s->runner.shutdown();
            delete s;
            return rc;
        }
    }
    for (int w = 0; w < s->opts.workers; w++) s->workers.emplace_back([s, w] { s->run(w); });
    *out = s;
    return MISFIT_OK;
}

extern "C" int misfit_server_destroy(misfit_server* s) {
    if (!s) return MISFIT_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    s->runner.shutdown();  // every request has finished: the kernel leaves at its next claim
    for (auto& kv : s->graphs) misfit_ctx_destroy(kv.second->ctx);
    delete s;
    return MISFIT_OK;
}

extern "C" const char* misfit_server_last_error(const misfit_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" int misfit_server_add_graph(misfit_server* s, int32_t n, int64_t m, const void* nindex, int nindex_is_64,
                                       const int32_t* nlist, const uint8_t* mis, int32_t* graph) {
    if (!s || !graph || !mis) return MISFIT_E_ARG;
    if (n < 1 || m < 0) return s->fail(MISFIT_E_ARG, "invalid graph size");
    const int64_t headroom = s->opts.compact_fraction > 0 ? compaction_headroom(n, m, s->opts.compact_fraction) : 0;
    // graphs that keep their compaction buffers hold their own headroom from
    // their first compaction on: it is charged with their resident bytes
    const int64_t need = graph_footprint(n, m) + (s->keep_compact ? headroom : 0);
    {
        // the graph's resident bytes, and the headroom of the largest
        // compaction any resident graph (or this one) may run, must fit
        std::lock_guard<std::mutex> lk(s->mu);
        const int64_t fr = device_free_bytes(s->device) - s->loading_bytes;
        const int64_t room = s->keep_compact ? 0 : std::max(headroom, s->headroom_max);
        if (fr - need - room < s->opts.hbm_reserve)
            return s->fail(MISFIT_E_NOMEM, "graph not admitted: needs " + std::to_string(need >> 20) + " MiB and " +
                                               std::to_string(room >> 20) + " MiB of compaction headroom, " +
                                               std::to_string(fr >> 20) + " MiB free, " +
                                               std::to_string(s->opts.hbm_reserve >> 20) + " MiB reserved");
        s->loading_bytes += need;
    }
    // the context is created and loaded without the lock: the workers go on
    auto g = std::make_unique<Graph>();
    int rc = misfit_ctx_create(s->device, &g->ctx);
    std::string em = "cannot create a context";
    if (rc == MISFIT_OK) {
        mf_ctx_set_keep_compact(g->ctx, s->keep_compact);
        rc = misfit_load_csr(g->ctx, n, m, nindex, nindex_is_64, nlist);
        if (rc == MISFIT_OK) rc = misfit_set_mis(g->ctx, mis);
        if (rc == MISFIT_OK) rc = misfit_set_mode(g->ctx, MISFIT_MODE_QUEUED);
        if (rc != MISFIT_OK) {
            em = misfit_last_error(g->ctx);
            misfit_ctx_destroy(g->ctx);
        }
    }
    if (rc == MISFIT_OK) mf_ctx_set_halting(g->ctx, s->opts.max_inflight == 1);
    std::lock_guard<std::mutex> lk(s->mu);
    s->loading_bytes -= need;
    if (rc != MISFIT_OK) return s->fail(rc, em);
    g->id = s->next_graph++;
    g->worker = g->id % s->opts.workers;
    g->n = n;
    g->base_m = m;
    s->headroom_max = std::max(s->headroom_max, headroom);
    *graph = g->id;
    s->graphs[g->id] = std::move(g);
    return MISFIT_OK;
}

// Waits (holding the lock on return) until no worker is inside graph id's
// context and none of its batches is in the serving kernel (whose writes
// the graph's own stream does not order); null when the graph is unknown.
static Graph* graph_idle(misfit_server* s, std::unique_lock<std::mutex>& lk, int32_t id) {
    while (true) {
        Graph* g = s->graph(id);
        if (!g || (!g->in_call && g->external == 0)) return g;
        s->cv_done.wait(lk);
    }
}

extern "C" int misfit_server_set_reserve(misfit_server* s, int64_t hbm_reserve) {
    if (!s || hbm_reserve < 0) return MISFIT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->opts.hbm_reserve = hbm_reserve;
    return MISFIT_OK;
}

extern "C" int misfit_server_remove_graph(misfit_server* s, int32_t graph) {
    if (!s) return MISFIT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (!graph_idle(s, lk, graph)) return s->fail(MISFIT_E_ARG, "unknown graph");
    auto it = s->graphs.find(graph);
    if (!it->second->pending.empty() || !it->second->inflight.empty() || it->second->compact_pending)
        return s->fail(MISFIT_E_STATE, "graph has requests in progress");
    misfit_ctx_destroy(it->second->ctx);
    s->graphs.erase(it);
    s->headroom_max = 0;
    if (s->opts.compact_fraction > 0)
        for (auto& kv : s->graphs)
            s->headroom_max = std::max(s->headroom_max,
                                       compaction_headroom(kv.second->n, kv.second->base_m, s->opts.compact_fraction));
    return MISFIT_OK;
}

extern "C" int misfit_server_submit(misfit_server* s, int32_t graph, const misfit_update* ups, int64_t count,
                                    uint64_t* request) {
    if (!s || count < 0 || (count && !ups) || !request) return MISFIT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->graphs.find(graph);
    if (it == s->graphs.end()) return s->fail(MISFIT_E_ARG, "unknown graph");
    auto r = std::make_unique<Request>();
    r->id = s->next_req++;
    r->graph = graph;
    r->ups.assign(ups, ups + count);
    r->t_submit = clk::now();
    r->st.graph = graph;
    r->st.updates = count;
    *request = r->id;
    it->second->pending.push_back(r.get());
    s->requests[r->id] = std::move(r);
    lk.unlock();
    s->cv_work.notify_all();  // the owning worker may be any of them
    return MISFIT_OK;
}

extern "C" int misfit_server_wait(misfit_server* s, uint64_t request, misfit_request_stats* st) {
    if (!s) return MISFIT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(request);
    if (it == s->requests.end()) return s->fail(MISFIT_E_ARG, "unknown or already collected request");
    Request* r = it->second.get();
    s->cv_done.wait(lk, [r] { return r->state == DONE; });
    if (st) *st = r->st;
    int rc = r->rc;
    if (rc) s->err = r->err;
    s->requests.erase(it);
    return rc;
}

// The state after every batch of the graph that has finished or is on the
// device (call after waiting for the graph's requests for a settled state).
extern "C" int misfit_server_get_mis(misfit_server* s, int32_t graph, uint8_t* mis) {
    if (!s || !mis) return MISFIT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    Graph* g = graph_idle(s, lk, graph);
    if (!g) return s->fail(MISFIT_E_ARG, "unknown graph");
    int rc = misfit_get_mis(g->ctx, mis);
    if (rc) s->err = misfit_last_error(g->ctx);
    return rc;
}

extern "C" int misfit_server_check(misfit_server* s, int32_t graph, int64_t* conflicts, int64_t* uncovered,
                                   int64_t* count_errors) {
    if (!s) return MISFIT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    Graph* g = graph_idle(s, lk, graph);
    if (!g) return s->fail(MISFIT_E_ARG, "unknown graph");
    int rc = misfit_check(g->ctx, conflicts, uncovered, count_errors);
    if (rc) s->err = misfit_last_error(g->ctx);
    return rc;
}
