// serve.hip -- the serving layer (host code only): many graphs held on one GPU, update
// batches taken as they arrive.
//
// The reference's misfit_async overlaps the host work of batch i+1 with the
// device work of batch i for one graph (src/misfit_async/misfit_async.cu:
// 484-701,827-876).  This generalises it to continuous batching:
//   * graphs are admitted by free HBM (their device footprint must fit in
//     the free memory minus a reserve), each in its own context and stream;
//   * requests (update batches) are queued per graph in arrival order; a
//     dispatcher thread keeps up to max_inflight batches of every graph
//     queued on the device, so batches of different graphs overlap on the
//     GPU and a request that arrives while others run starts as soon as its
//     graph has a free slot; a batch whose buffers may not fit in free HBM
//     is refused (MISFIT_E_NOMEM) instead of being started;
//   * a graph whose overlay passes compact_fraction of its base entries is
//     compacted once its in-flight batches have drained (its later requests
//     wait); every request reports how many compactions of its graph ran
//     before it started, which places each compaction exactly;
//   * every request reports its queueing time, its latency from submission
//     to result and its device time.
// One dispatcher thread owns every context call for batches; the public
// calls take the server lock, so any number of client threads may submit.
#include <hip/hip_runtime.h>

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
    ReqState state = PENDING;
    int rc = MISFIT_OK;
    std::string err;
    misfit_request_stats st{};
};

struct Graph {
    int32_t id = -1;
    misfit_ctx* ctx = nullptr;
    int32_t n = 0;
    int64_t base_m = 0;
    std::deque<Request*> pending, inflight;
    bool compact_pending = false;
    int64_t compactions = 0;
};

// Device bytes a graph needs beyond its CSR: per-vertex state, the check
// scratch, the per-batch lists and a first overlay pool and hash.
int64_t graph_footprint(int32_t n, int64_t m) {
    return 8 * ((int64_t)n + 1) + 4 * m + 256 * (int64_t)n + (int64_t(64) << 20);
}

// Device bytes one batch of `count` updates may add (batch buffers doubled
// to the new capacity, overlay pool and inserted-pair hash growth).
int64_t batch_footprint(int64_t count) { return 1024 * count; }
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
    bool stop = false;
    std::thread dispatcher;
    std::string err;

    int fail(int rc, const std::string& m) {
        err = m;
        return rc;
    }
    bool busy() const {
        for (auto& kv : graphs)
            if (!kv.second->pending.empty() || !kv.second->inflight.empty()) return true;
        return false;
    }
    bool anything_inflight() const {
        for (auto& kv : graphs)
            if (!kv.second->inflight.empty()) return true;
        return false;
    }
    void finish(Request* r, int rc, const std::string& msg) {
        r->rc = rc;
        r->err = msg;
        r->state = DONE;
        r->st.latency_ms = ms_between(r->t_submit, clk::now());
        cv_done.notify_all();
    }
    int64_t free_bytes() {
        size_t fr = 0, tot = 0;
        hipSetDevice(device);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
        return (int64_t)fr;
    }
    void run();
};

// One pass over the graphs: collect finished batches, compact where due,
// start queued requests.  Returns whether anything happened.  Holds mu.
static bool dispatch_once(misfit_server* s) {
    bool progress = false;
    for (auto& kv : s->graphs) {
        Graph* g = kv.second.get();
        while (!g->inflight.empty()) {
            Request* r = g->inflight.front();
            int p = misfit_poll(g->ctx, r->ticket);
            if (p == 0) break;
            g->inflight.pop_front();
            progress = true;
            misfit_batch_stats bs;
            int rc = p < 0 ? p : misfit_wait(g->ctx, r->ticket, &bs);
            if (rc != MISFIT_OK) {
                s->finish(r, rc, misfit_last_error(g->ctx));
                continue;
            }
            r->st.batch = bs;
            r->st.device_ms = bs.device_ms;
            if (s->opts.compact_fraction > 0 && (double)bs.overlay_entries > s->opts.compact_fraction * (double)g->base_m)
                g->compact_pending = true;
            s->finish(r, MISFIT_OK, "");
        }
        if (g->compact_pending && g->inflight.empty()) {
            int64_t new_m = 0;
            int rc = misfit_compact(g->ctx, &new_m);
            g->compact_pending = false;
            if (rc == MISFIT_OK) {
                g->base_m = new_m;
                g->compactions++;
            }
            progress = true;
        }
        while (!g->compact_pending && (int)g->inflight.size() < s->opts.max_inflight && !g->pending.empty()) {
            Request* r = g->pending.front();
            g->pending.pop_front();
            progress = true;
            if (s->free_bytes() - batch_footprint((int64_t)r->ups.size()) < s->opts.hbm_reserve) {
                s->finish(r, MISFIT_E_NOMEM, "batch not admitted: not enough free device memory");
                continue;
            }
            int rc = misfit_submit_batch(g->ctx, r->ups.data(), (int64_t)r->ups.size(), &r->ticket);
            if (rc != MISFIT_OK) {
                s->finish(r, rc, misfit_last_error(g->ctx));
                continue;
            }
            r->t_enqueue = clk::now();
            r->st.compactions_before = g->compactions;
            r->st.queue_ms = ms_between(r->t_submit, r->t_enqueue);
            r->state = INFLIGHT;
            g->inflight.push_back(r);
        }
    }
    return progress;
}

void misfit_server::run() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (stop && !busy()) break;
        if (dispatch_once(this)) continue;
        if (!anything_inflight()) {
            cv_work.wait(lk);  // until a submit or the stop request
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
    s->dispatcher = std::thread([s] { s->run(); });
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
    s->dispatcher.join();
    for (auto& kv : s->graphs) misfit_ctx_destroy(kv.second->ctx);
    delete s;
    return MISFIT_OK;
}

extern "C" const char* misfit_server_last_error(const misfit_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" int misfit_server_add_graph(misfit_server* s, int32_t n, int64_t m, const void* nindex, int nindex_is_64,
                                       const int32_t* nlist, const uint8_t* mis, int32_t* graph) {
    if (!s || !graph || !mis) return MISFIT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t need = graph_footprint(n, m);
    int64_t fr = s->free_bytes();
    if (fr - need < s->opts.hbm_reserve)
        return s->fail(MISFIT_E_NOMEM, "graph not admitted: needs " + std::to_string(need >> 20) + " MiB, " +
                                           std::to_string(fr >> 20) + " MiB free, " +
                                           std::to_string(s->opts.hbm_reserve >> 20) + " MiB reserved");
    auto g = std::make_unique<Graph>();
    int rc = misfit_ctx_create(s->device, &g->ctx);
    if (rc) return s->fail(rc, "cannot create a context");
    rc = misfit_load_csr(g->ctx, n, m, nindex, nindex_is_64, nlist);
    if (rc == MISFIT_OK) rc = misfit_set_mis(g->ctx, mis);
    if (rc == MISFIT_OK) rc = misfit_set_mode(g->ctx, MISFIT_MODE_QUEUED);
    if (rc != MISFIT_OK) {
        std::string m2 = misfit_last_error(g->ctx);
        misfit_ctx_destroy(g->ctx);
        return s->fail(rc, m2);
    }
    g->id = s->next_graph++;
    g->n = n;
    g->base_m = m;
    *graph = g->id;
    s->graphs[g->id] = std::move(g);
    return MISFIT_OK;
}

extern "C" int misfit_server_remove_graph(misfit_server* s, int32_t graph) {
    if (!s) return MISFIT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->graphs.find(graph);
    if (it == s->graphs.end()) return s->fail(MISFIT_E_ARG, "unknown graph");
    if (!it->second->pending.empty() || !it->second->inflight.empty())
        return s->fail(MISFIT_E_STATE, "graph has requests in progress");
    misfit_ctx_destroy(it->second->ctx);
    s->graphs.erase(it);
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
    s->cv_work.notify_one();
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
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->graphs.find(graph);
    if (it == s->graphs.end()) return s->fail(MISFIT_E_ARG, "unknown graph");
    int rc = misfit_get_mis(it->second->ctx, mis);
    if (rc) s->err = misfit_last_error(it->second->ctx);
    return rc;
}

extern "C" int misfit_server_check(misfit_server* s, int32_t graph, int64_t* conflicts, int64_t* uncovered,
                                   int64_t* count_errors) {
    if (!s) return MISFIT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->graphs.find(graph);
    if (it == s->graphs.end()) return s->fail(MISFIT_E_ARG, "unknown graph");
    int rc = misfit_check(it->second->ctx, conflicts, uncovered, count_errors);
    if (rc) s->err = misfit_last_error(it->second->ctx);
    return rc;
}
