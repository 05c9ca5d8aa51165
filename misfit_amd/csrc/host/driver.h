// driver.h -- shared pieces of the drop-in binaries `misfit` and
// `misfit_async`: argument handling, MIS / batch-folder input, the final MIS
// file, end-of-run compaction and the report.  The engine itself is reached
// only through the C ABI (include/misfit.h).
//
// Process contract followed (reference = anonymous-artifacts/MISFIT):
//   argv: <graph.egr> <initial_mis.txt> <batch_folder> <num_batches> <gpu_id>
//         (core/cli_setup.h:28-41, src/misfit/misfit.cu:737-743)
//   MIS file: one 1-indexed id per line; out-of-range ids ignored and the
//         initial cardinality counts in-range lines (src/misfit/misfit.cu:778-791)
//   batches: every regular file of the folder, ordered by the integer value
//         of its stem; the first <num_batches> are used; each line "u v"
//         (1-indexed); out-of-range pairs and self-loops dropped; insertion
//         flag = not adjacent in the graph as loaded, fixed for all batches
//         before any is replayed (core/cli_setup.h:49-82, misfit.cu:802-828)
//   output: datasets/Final_MISs/<stem>_flash_final_mis.txt under the current
//         directory (misfit.cu:892-907); the input .egr rewritten in the wide
//         form after compaction unless MISFIT_SKIP_COMPACTION=1
//         (misfit.cu:909-926, core/flash_compactor.h:40-104).
// Additions (environment, never changing the 5-argument contract):
//   MISFIT_RULE=reference   literal reference rule without completion
//   MISFIT_MODE=plain|queued   how batches are driven (default queued)
//   MISFIT_CHECK=0          skip the full check after every batch
#pragma once
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/misfit.h"

namespace mfd {

using clk = std::chrono::high_resolution_clock;

inline double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

[[noreturn]] inline void die(const std::string& msg) {
    std::cerr << "ERROR: " << msg << std::endl;
    std::exit(1);
}

inline void check(misfit_ctx* ctx, int rc, const char* what) {
    if (rc != MISFIT_OK) die(std::string(what) + " failed (" + std::to_string(rc) + "): " + misfit_last_error(ctx));
}

// istream >> int on a line: skip blanks, optional sign, digits; stops at the
// first non-digit; fails without digits or on overflow.
inline const char* parse_int(const char* p, int32_t* out) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = (*p++ == '-');
    if (*p < '0' || *p > '9') return nullptr;
    int64_t v = 0;
    while (*p >= '0' && *p <= '9') {
        v = v * 10 + (*p - '0');
        if (v > (int64_t)INT32_MAX + 1) return nullptr;
        p++;
    }
    if (neg) v = -v;
    if (v > INT32_MAX || v < INT32_MIN) return nullptr;
    *out = (int32_t)v;
    return p;
}

struct Args {
    std::string egr, mis, batch_dir;
    int num_batches = 0;
    int gpu = 0;
};

inline Args parse_args(int argc, char** argv) {
    if (argc < 6) {
        std::cerr << "Usage: " << argv[0] << " <graph.egr> <initial_mis.txt> <batch_folder> <num_batches> <gpu_device_id>"
                  << std::endl;
        std::exit(1);
    }
    Args a;
    a.egr = argv[1];
    a.mis = argv[2];
    a.batch_dir = argv[3];
    int32_t v;
    if (!parse_int(argv[4], &v)) die(std::string("num_batches is not an integer: ") + argv[4]);
    a.num_batches = v;
    if (!parse_int(argv[5], &v)) die(std::string("gpu_device_id is not an integer: ") + argv[5]);
    a.gpu = v;
    if (a.egr.size() < 4 || a.egr.substr(a.egr.size() - 4) != ".egr")
        std::cerr << "[Warning] " << a.egr << " does not end in .egr — this binary expects a pre-built binary CSR file."
                  << std::endl;
    return a;
}

inline std::vector<uint8_t> read_mis(const std::string& path, int32_t n, long long* card) {
    std::vector<uint8_t> mis((size_t)n, 0);
    *card = 0;
    std::ifstream f(path);
    if (!f.is_open()) return mis;  // the reference starts from an empty set
    std::string line;
    long long lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        int32_t id;
        if (!parse_int(line.c_str(), &id)) die(path + ":" + std::to_string(lineno) + ": not an integer");
        id -= 1;
        if (id >= 0 && id < n) {
            mis[(size_t)id] = 1;
            (*card)++;
        }
    }
    return mis;
}

inline std::vector<std::filesystem::path> batch_files(const std::string& dir) {
    std::vector<std::pair<long long, std::filesystem::path>> v;
    std::error_code ec;
    for (const auto& e : std::filesystem::directory_iterator(dir, ec)) {
        if (!e.is_regular_file()) continue;
        std::string stem = e.path().stem().string();
        int32_t k;
        if (!parse_int(stem.c_str(), &k)) die("batch file name is not numbered: " + e.path().string());
        v.push_back({k, e.path()});
    }
    if (ec) die("cannot list batch folder " + dir + ": " + ec.message());
    std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<std::filesystem::path> out;
    for (auto& p : v) out.push_back(p.second);
    return out;
}

struct Batch {
    std::vector<misfit_update> ups;
};

// Parses up to num_batches files; flags against the graph as loaded (the
// context holds exactly the loaded graph at this point).
inline std::vector<Batch> read_batches(misfit_ctx* ctx, const std::string& dir, int num_batches, int32_t n) {
    std::vector<Batch> out;
    for (const auto& path : batch_files(dir)) {
        if ((int)out.size() >= num_batches) break;
        std::ifstream f(path);
        std::string line;
        std::vector<int32_t> us, vs;
        while (std::getline(f, line)) {
            int32_t a, b;
            const char* p = parse_int(line.c_str(), &a);
            if (!p) continue;
            if (!parse_int(p, &b)) continue;
            a--;
            b--;
            if (a >= 0 && a < n && b >= 0 && b < n && a != b) {
                us.push_back(a);
                vs.push_back(b);
            }
        }
        std::vector<uint8_t> ins(us.size());
        if (!us.empty()) check(ctx, misfit_flags(ctx, us.data(), vs.data(), (int64_t)us.size(), ins.data()), "flags");
        Batch bt;
        bt.ups.resize(us.size());
        for (size_t i = 0; i < us.size(); i++) {
            bt.ups[i].u = us[i];
            bt.ups[i].v = vs[i];
            bt.ups[i].is_insertion = ins[i];
            memset(bt.ups[i].pad, 0, 3);
        }
        out.push_back(std::move(bt));
    }
    return out;
}

inline bool env_is(const char* name, const char* val) {
    const char* e = getenv(name);
    return e && std::string(e) == val;
}

struct Setup {
    misfit_ctx* ctx = nullptr;
    int32_t n = 0;
    int64_t m = 0;
    int wide = 0;
    long long initial_card = 0;
    std::vector<Batch> batches;
    std::string device_name;
    int64_t mem_total = 0, mem_base_used = 0, peak_used = 0;
    bool check_each = true;
};

inline void sample_memory(Setup& S) {
    int64_t tot = 0, fr = 0;
    if (misfit_device_info(S.ctx, nullptr, 0, &tot, &fr) == MISFIT_OK) {
        int64_t used = tot - fr - S.mem_base_used;
        if (used > S.peak_used) S.peak_used = used;
    }
}

inline Setup setup(const Args& a) {
    Setup S;
    if (misfit_ctx_create(a.gpu, &S.ctx) != MISFIT_OK) die("cannot open GPU " + std::to_string(a.gpu));
    char name[256] = {0};
    int64_t tot = 0, fr = 0;
    check(S.ctx, misfit_device_info(S.ctx, name, sizeof(name), &tot, &fr), "device info");
    S.device_name = name;
    S.mem_total = tot;
    S.mem_base_used = tot - fr;
    misfit_egr g;
    char err[512];
    if (misfit_egr_open(a.egr.c_str(), &g, err, sizeof(err)) != MISFIT_OK) die(std::string("FlashGraph: ") + err);
    S.n = g.n;
    S.m = g.m;
    S.wide = g.wide;
    check(S.ctx, misfit_load_csr(S.ctx, g.n, g.m, g.nindex, g.wide, g.nlist), "load graph");
    misfit_egr_close(&g);  // the base now lives in device memory
    sample_memory(S);
    std::vector<uint8_t> mis = read_mis(a.mis, S.n, &S.initial_card);
    check(S.ctx, misfit_set_mis(S.ctx, mis.data()), "set MIS");
    check(S.ctx, misfit_set_rule(S.ctx, env_is("MISFIT_RULE", "reference") ? 0 : 1), "set rule");
    if (env_is("MISFIT_MODE", "plain")) check(S.ctx, misfit_set_mode(S.ctx, MISFIT_MODE_PLAIN), "set mode");
    S.check_each = !env_is("MISFIT_CHECK", "0");
    S.batches = read_batches(S.ctx, a.batch_dir, a.num_batches, S.n);
    return S;
}

inline double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double r = q * (double)(v.size() - 1);
    size_t lo = (size_t)std::floor(r), hi = (size_t)std::ceil(r);
    return v[lo] + (v[hi] - v[lo]) * (r - (double)lo);
}

// The reference's per-batch breakdown, line for line (misfit.cu:853-870,
// misfit_async.cu:853-870), with this engine's figures: 1 = the host's
// staging and launching of the batch; 3 = the prologue on the device
// (overlay update, may-uncover set, admissible index); 4 = the exact rule's
// speculative rounds, the tail and the completion; 7 = the rest of the
// host's time for the batch; 8 = speculative rounds.  Stages this engine
// does not have (the first kernel reads the batch from pinned memory,
// results come back through pinned memory, the overlay lives on the device,
// nothing is clustered) print 0.000.  This engine's own detail follows on lines of
// its own after "Finished MIS".
struct BatchTimes {
    double total_ms = 0;   // from the start of its submission to the end of its collection
    double host_ms = 0;    // the submission call (staging and launches)
};

inline void print_batch_detail(int i, const misfit_batch_stats& st, const BatchTimes& t, bool async, bool checked,
                               int64_t conf, int64_t unc) {
    const double z = 0.0;
    const double gpu_rounds = std::max(0.0, st.device_ms - st.prologue_ms);
    const double overhead = std::max(0.0, t.total_ms - t.host_ms - st.device_ms);
    std::cout << (async ? "=== BATCH TIMING BREAKDOWN (algorithm-specific detail) ===" : "=== BATCH TIMING BREAKDOWN ===")
              << std::endl;
    std::cout << std::fixed << std::setprecision(3);
    std::cout << "Total Batch Time: " << t.total_ms << " ms\n\n";
    std::cout << "Component Breakdown:" << std::endl;
    std::cout << "  1. CPU Conversion:        " << t.host_ms << " ms (CPU)" << std::endl;
    std::cout << "  2. Host->Device Transfer: " << z << " ms (GPU)" << std::endl;
    std::cout << "  3. Cluster Processing:    " << st.prologue_ms << " ms (GPU)" << std::endl;
    std::cout << "       - Extract vertices: " << z << " ms" << std::endl;
    std::cout << "       - Compute neighb.:  " << z << " ms" << std::endl;
    std::cout << "       - Vertex cluster.:  " << z << " ms" << std::endl;
    std::cout << "       - Assign edges:     " << z << " ms" << std::endl;
    std::cout << "       - Clear assigns.:   " << z << " ms" << std::endl;
    std::cout << "  4. GPU Process:           " << gpu_rounds << " ms (GPU)" << std::endl;
    std::cout << "  5. Device->Host Transfer: " << z << " ms (GPU)" << std::endl;
    std::cout << "  6. CPU Graph Update:      " << z << " ms (CPU)" << std::endl;
    std::cout << "  7. CPU Overhead:          " << overhead << " ms (CPU)" << std::endl;
    std::cout << "  8. Cluster count:         " << st.rounds << std::endl;
    std::cout << "Finished MIS for batch " << i << " -> slot " << (i % 2) << std::endl;
    std::cout << "  Engine detail: device " << st.device_ms << " ms, rounds " << st.rounds << " ("
              << st.evaluations << " edge evaluations), evictions " << st.evictions << ", admissions "
              << (st.admit_ins + st.admit_del) << " (+" << st.completion_adds << " by completion), may-uncover "
              << st.maybe_uncovered << " (" << st.admissible << " admissible), overlay operations " << st.delta_ops
              << std::endl;
    if (checked) std::cout << "  Check: conflicts=" << conf << " uncovered=" << unc << std::endl;
}

inline void finish(Setup& S, const Args& a, const std::vector<double>& lat_ms, double time_cum, double setup_ms,
                   bool print_summary, double sum_prologue_ms, double sum_rounds_ms, const std::string& note) {
    misfit_sync(S.ctx);
    auto d0 = clk::now();
    std::vector<uint8_t> mis((size_t)S.n);
    check(S.ctx, misfit_get_mis(S.ctx, mis.data()), "get MIS");
    double d2h_ms = ms_since(d0);
    long long card = 0;
    for (uint8_t x : mis) card += x;
    {
        std::string out_dir = "datasets/Final_MISs";
        std::error_code ec;
        std::filesystem::create_directories(out_dir, ec);
        std::string stem = std::filesystem::path(a.egr).stem().string();
        std::string path = out_dir + "/" + stem + "_flash_final_mis.txt";
        std::ofstream f(path);
        if (f.is_open()) {
            for (int32_t i = 0; i < S.n; i++)
                if (mis[(size_t)i]) f << (i + 1) << "\n";
            f.close();
            std::cout << "Saved final MIS to " << path << std::endl;
        } else {
            std::cerr << "Warning: Could not open file to save final MIS: " << path << std::endl;
        }
    }
    double compact_ms = 0.0;
    if (!env_is("MISFIT_SKIP_COMPACTION", "1")) {
        auto c0 = clk::now();
        int64_t new_m = 0;
        check(S.ctx, misfit_compact(S.ctx, &new_m), "compaction");
        std::vector<int64_t> ni((size_t)S.n + 1);
        std::vector<int32_t> nl((size_t)std::max<int64_t>(new_m, 1));
        check(S.ctx, misfit_export_csr(S.ctx, ni.data(), nl.data()), "export");
        char err[512];
        if (misfit_write_egr(a.egr.c_str(), S.n, new_m, ni.data(), nl.data(), err, sizeof(err)) != MISFIT_OK)
            die(std::string("flash_compactor: ") + err);
        compact_ms = ms_since(c0);
    } else {
        std::cout << "\n[MISFIT_SKIP_COMPACTION=1] Skipping end-of-cycle compaction — on-disk .egr left untouched."
                  << std::endl;
    }
    int processed = (int)lat_ms.size();
    int denom = processed ? processed : 1;
    long long updates = 0;
    for (auto& b : S.batches) updates += (long long)b.ups.size();
    std::cout << "\n" << std::string(80, '=') << std::endl;
    std::cout << "=== FINAL RESULTS ===" << std::endl;
    std::cout << std::string(80, '=') << "\n" << std::endl;
    std::cout << "Graph Statistics:" << std::endl;
    std::cout << "  Initial MIS cardinality: " << S.initial_card << std::endl;
    std::cout << "  Final MIS cardinality:   " << card << std::endl;
    std::cout << "  Cardinality change:      " << std::llabs(S.initial_card - card) << std::endl;
    std::cout << "  Number of batches:       " << denom << "\n" << std::endl;
    std::cout << "Overall Timing Results:" << std::endl;
    std::cout << std::fixed << std::setprecision(3);
    std::cout << "  Total processing time:   " << time_cum << " milliseconds" << std::endl;
    std::cout << "  Average time per batch:  " << time_cum / denom << " milliseconds" << std::endl;
    std::cout << "  Final D2H Transfer Time: " << d2h_ms << " milliseconds" << std::endl;
    std::cout << std::setprecision(1);
    std::cout << "  Processing rate:         " << (time_cum > 0 ? denom / (time_cum / 1000.0) : 0.0)
              << " batches/second\n" << std::endl;
    std::cout << "Average Component Times:" << std::endl;
    std::cout << std::fixed << std::setprecision(3);
    std::cout << "  Average Cluster Processing: " << (sum_prologue_ms / denom) << " ms" << std::endl;
    std::cout << "  Average TVB Time:           " << (sum_rounds_ms / denom) << " ms" << std::endl;
    std::cout << "\nEnd-of-cycle compaction time (NOT counted in per-batch numbers above): " << compact_ms << " ms"
              << std::endl;
    sample_memory(S);
    double peak_mb = (double)S.peak_used / (1024.0 * 1024.0);
    double tot_mb = (double)S.mem_total / (1024.0 * 1024.0);
    std::cout << "Peak program GPU memory: " << std::setprecision(2) << peak_mb << " MB (" << std::setprecision(1)
              << (tot_mb > 0 ? peak_mb / tot_mb * 100.0 : 0.0) << "% of " << std::setprecision(2) << tot_mb
              << " MB total)" << std::endl;
    std::cout << std::string(80, '=') << std::endl;
    if (print_summary) {
        double e2e = setup_ms + time_cum + compact_ms;
        std::cout << "\n=== FINAL SUMMARY ===" << std::endl;
        std::cout << "Setup time:      " << std::fixed << std::setprecision(3) << setup_ms << " ms" << std::endl;
        std::cout << "Processing time: " << std::fixed << std::setprecision(3) << time_cum << " ms" << std::endl;
        std::cout << "End-to-end time: " << std::fixed << std::setprecision(3) << e2e << " ms" << std::endl;
        std::cout << "Initial MIS size: " << S.initial_card << std::endl;
        std::cout << "Final MIS size:   " << card << std::endl;
    }
    // this engine's additions, after everything the reference prints
    std::cout << "\nEngine summary:" << std::endl;
    std::cout << std::setprecision(1);
    std::cout << "  Edge updates/second:     " << (time_cum > 0 ? (double)updates / (time_cum / 1000.0) : 0.0)
              << std::endl;
    std::cout << std::setprecision(3);
    std::cout << "  Batch latency p50/p99:   " << percentile(lat_ms, 0.50) << " / " << percentile(lat_ms, 0.99)
              << " ms" << std::endl;
    std::cout << "  GPU: " << S.device_name << std::endl;
    if (!note.empty()) std::cout << note << std::endl;
    misfit_ctx_destroy(S.ctx);
}

}  // namespace mfd
