// gpu_vertex_clustering_graph.h -- the reference's engine class over the
// MI355X C ABI (include/misfit.h), for driver code written against the
// reference's GPUVertexClusteringGraph (src/misfit/misfit.cu:360-693 and
// src/misfit_async/misfit_async.cu:360-701).  Header only, C++17; link with
// -lmisfit_amd.
//
// The member names and argument order are the reference's.  What differs,
// and why:
//   * the constructor keeps (total_nodes, max_csr_elements); the engine
//     stages no packed neighbourhood rows, so max_csr_elements is accepted
//     and unused.  The base graph is streamed into device memory once, by
//     loadBase, where the reference reads it through its HybridGraph at every
//     batch;
//   * the HybridGraph* argument of the batch calls is not read: the overlay
//     lives on the device and each batch updates it there (the reference's
//     caller applies the batch to its host overlay first,
//     src/misfit/misfit.cu:704-714).  Callers keep passing it;
//   * hop_limit must be 2, the reference driver's only value
//     (src/misfit/misfit.cu:831); the rule's rows are two hops;
//   * errors throw std::runtime_error (the reference's CUDA_CHECK exits).
//
// Edge is the caller's edge record with members source, destination and
// isInsertion (the reference's DeviceEdge); Timings is the caller's
// per-batch timing record with the reference DetailedTimings' members.  A
// batch fills: host_to_device_time (the host hand-over), cpu_graph_update_time
// (0: the overlay update runs on the device), compute_neighborhoods_time (the
// device prologue: overlay update, may-uncover set, admissible index),
// tvb_process_time (the repair rounds and the completion), clusters_count (the
// speculative rounds), total_batch_time (launch to collect, on the host).
#pragma once

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/misfit.h"

template <class Edge, class Timings>
class GPUVertexClusteringGraphT {
  public:
    // What launchBatch hands to syncAndCollect (src/misfit_async/misfit_async.cu:365).
    struct LaunchedBatch {
        uint64_t ticket = 0;
        size_t edges = 0;
        double host_to_device_ms = 0.0;
        std::chrono::steady_clock::time_point start;
    };

    GPUVertexClusteringGraphT(int total_nodes, long long max_csr_elements, int device = 0)
        : total_nodes_(total_nodes) {
        (void)max_csr_elements;
        ok(misfit_ctx_create(device, &ctx_), "misfit_ctx_create");
    }
    ~GPUVertexClusteringGraphT() {
        if (ctx_) misfit_ctx_destroy(ctx_);
    }
    GPUVertexClusteringGraphT(const GPUVertexClusteringGraphT&) = delete;
    GPUVertexClusteringGraphT& operator=(const GPUVertexClusteringGraphT&) = delete;

    // The base graph, once, before the first batch (a view from misfit_egr_open).
    void loadBase(const misfit_egr& g) {
        if (g.n != total_nodes_) throw std::runtime_error("misfit: graph size differs from total_nodes");
        ok(misfit_load_csr(ctx_, g.n, g.m, g.nindex, g.wide, g.nlist), "misfit_load_csr");
    }

    void setMISFromHost(const std::vector<bool>& mis) {  // src/misfit/misfit.cu:448-453
        std::vector<uint8_t> b(mis.begin(), mis.end());
        b.resize((size_t)total_nodes_, 0);
        ok(misfit_set_mis(ctx_, b.data()), "misfit_set_mis");
    }

    std::vector<bool> getMISToHost() {  // src/misfit/misfit.cu:455-462
        std::vector<uint8_t> b((size_t)total_nodes_);
        ok(misfit_get_mis(ctx_, b.data()), "misfit_get_mis");
        return std::vector<bool>(b.begin(), b.end());
    }

    void synchronizeChannels() { ok(misfit_sync(ctx_), "misfit_sync"); }  // src/misfit/misfit.cu:464-466

    // src/misfit/misfit.cu:468-730: the whole batch, waited for.
    template <class HostGraph>
    void processBatchWithVertexClustering(const std::vector<Edge>& batch, int hop_limit, HostGraph* cpu_graph,
                                          Timings& timings) {
        LaunchedBatch lb = launchBatch(batch, hop_limit, cpu_graph, timings);
        syncAndCollect(lb, timings);
    }

    // src/misfit_async/misfit_async.cu:484-675: queues the batch and returns
    // at once (the update records are copied before it returns).
    template <class HostGraph>
    LaunchedBatch launchBatch(const std::vector<Edge>& batch, int hop_limit, HostGraph* cpu_graph, Timings& timings) {
        (void)cpu_graph;
        if (hop_limit != 2) throw std::runtime_error("misfit: only hop_limit 2 (the rule's two-hop rows)");
        timings.reset();
        LaunchedBatch lb;
        lb.start = std::chrono::steady_clock::now();
        std::vector<misfit_update> ups(batch.size());
        for (size_t i = 0; i < batch.size(); i++) {
            ups[i].u = batch[i].source;
            ups[i].v = batch[i].destination;
            ups[i].is_insertion = batch[i].isInsertion ? 1 : 0;
        }
        ok(misfit_submit_batch(ctx_, ups.data(), (int64_t)ups.size(), &lb.ticket), "misfit_submit_batch");
        lb.edges = batch.size();
        lb.host_to_device_ms = ms_since(lb.start);
        return lb;
    }

    // src/misfit_async/misfit_async.cu:677-701.
    void syncAndCollect(LaunchedBatch& launched, Timings& timings) {
        misfit_batch_stats st;
        ok(misfit_wait(ctx_, launched.ticket, &st), "misfit_wait");
        timings.host_to_device_time = (float)launched.host_to_device_ms;
        timings.cpu_graph_update_time = 0.0f;
        timings.compute_neighborhoods_time = (float)st.prologue_ms;
        timings.tvb_process_time = (float)(st.device_ms - st.prologue_ms);
        timings.clusters_count = (int)st.rounds;
        timings.calculateAggregates();
        timings.total_batch_time = (float)ms_since(launched.start);
        last_ = st;
    }

    const misfit_batch_stats& lastStats() const { return last_; }
    misfit_ctx* ctx() { return ctx_; }

  private:
    static double ms_since(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    }
    void ok(int rc, const char* what) {
        if (rc != MISFIT_OK)
            throw std::runtime_error(std::string("misfit: ") + what + ": " + (ctx_ ? misfit_last_error(ctx_) : ""));
    }
    misfit_ctx* ctx_ = nullptr;
    int total_nodes_ = 0;
    misfit_batch_stats last_{};
};
