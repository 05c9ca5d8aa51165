// Drives the bundled sample (graph, initial MIS, five batch files) through
// the reference-shaped engine class, once batch by batch with
// processBatchWithVertexClustering and once pipelined with launchBatch /
// syncAndCollect, and prints the final MIS of each run (1-indexed ids, one
// line per run).  Usage: shim_sample <graph.egr> <mis.txt> <batch_dir> <gpu>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../misfit_amd/shim/gpu_vertex_clustering_graph.h"

// The caller's records, with the members the engine class reads and fills
// (the reference driver's DeviceEdge and DetailedTimings).
struct Edge {
    int source, destination;
    bool isInsertion;
};
struct Timings {
    float host_to_device_time, cpu_graph_update_time, compute_neighborhoods_time, tvb_process_time;
    float cluster_processing_time, total_gpu_kernel_time, transfer_time, total_cpu_time, total_gpu_time;
    float total_batch_time;
    int clusters_count;
    void reset() { *this = Timings{}; }
    void calculateAggregates() {
        cluster_processing_time = compute_neighborhoods_time;
        total_gpu_kernel_time = cluster_processing_time + tvb_process_time;
        transfer_time = host_to_device_time;
        total_cpu_time = transfer_time + cpu_graph_update_time;
        total_gpu_time = total_gpu_kernel_time + transfer_time;
    }
};
struct NoHostGraph {};  // the overlay lives on the device

using Engine = GPUVertexClusteringGraphT<Edge, Timings>;

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s <graph.egr> <mis.txt> <batch_dir> <gpu>\n", argv[0]);
        return 1;
    }
    misfit_egr g;
    char err[256];
    if (misfit_egr_open(argv[1], &g, err, sizeof err) != MISFIT_OK) {
        fprintf(stderr, "%s\n", err);
        return 1;
    }
    std::vector<bool> mis0((size_t)g.n, false);
    {
        std::ifstream f(argv[2]);
        long long id;
        while (f >> id)
            if (id >= 1 && id <= g.n) mis0[(size_t)(id - 1)] = true;
    }
    std::vector<std::filesystem::path> files;
    for (auto& e : std::filesystem::directory_iterator(argv[3]))
        if (e.is_regular_file()) files.push_back(e.path());
    std::sort(files.begin(), files.end(), [](const auto& a, const auto& b) {
        return std::stoi(a.stem().string()) < std::stoi(b.stem().string());
    });
    const int gpu = atoi(argv[4]);
    try {
        std::vector<std::vector<Edge>> batches;
        {
            // insertion flags against the graph as loaded (core/cli_setup.h:77)
            Engine flags(g.n, 0, gpu);
            flags.loadBase(g);
            for (auto& p : files) {
                std::ifstream f(p);
                std::string line;
                std::vector<int32_t> u, v;
                while (std::getline(f, line)) {
                    std::istringstream is(line);
                    long long a, b;
                    if (!(is >> a >> b)) continue;
                    a -= 1;
                    b -= 1;
                    if (a < 0 || b < 0 || a >= g.n || b >= g.n || a == b) continue;
                    u.push_back((int32_t)a);
                    v.push_back((int32_t)b);
                }
                std::vector<uint8_t> ins(u.size());
                if (!u.empty() && misfit_flags(flags.ctx(), u.data(), v.data(), (int64_t)u.size(), ins.data()) != 0)
                    throw std::runtime_error("misfit_flags failed");
                std::vector<Edge> es(u.size());
                for (size_t i = 0; i < u.size(); i++) es[i] = Edge{u[i], v[i], ins[i] != 0};
                batches.push_back(std::move(es));
            }
        }
        NoHostGraph host;
        for (int run = 0; run < 2; run++) {
            Engine eng(g.n, 1000000, gpu);
            eng.loadBase(g);
            eng.setMISFromHost(mis0);
            std::vector<Timings> t(batches.size());
            if (run == 0) {
                for (size_t b = 0; b < batches.size(); b++)
                    eng.processBatchWithVertexClustering(batches[b], 2, &host, t[b]);
            } else {  // two in flight, as the reference's pipelined main loop
                Engine::LaunchedBatch cur = eng.launchBatch(batches[0], 2, &host, t[0]);
                for (size_t b = 0; b < batches.size(); b++) {
                    Engine::LaunchedBatch next;
                    if (b + 1 < batches.size()) next = eng.launchBatch(batches[b + 1], 2, &host, t[b + 1]);
                    eng.syncAndCollect(cur, t[b]);
                    cur = next;
                }
            }
            eng.synchronizeChannels();
            std::vector<bool> mis = eng.getMISToHost();
            for (int x = 0; x < g.n; x++)
                if (mis[(size_t)x]) printf("%d ", x + 1);
            printf("\n");
            for (size_t b = 0; b < t.size(); b++)
                if (t[b].total_batch_time <= 0.0f) throw std::runtime_error("batch time not filled");
        }
        try {
            Engine bad(g.n, 0, gpu);
            bad.loadBase(g);
            Timings tb;
            bad.processBatchWithVertexClustering(batches[0], 3, &host, tb);
            fprintf(stderr, "hop_limit 3 was accepted\n");
            return 1;
        } catch (const std::runtime_error&) {
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        misfit_egr_close(&g);
        return 1;
    }
    misfit_egr_close(&g);
    return 0;
}
