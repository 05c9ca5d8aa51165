// misfit_egr -- graph-file tool.
//   misfit_egr convert <input.mtx> <output.egr>
//        what tools/efficient_converter.py does (wide form, file row order)
//   misfit_egr check <graph.egr>
//        structure report; exit status 1 when the file is not a valid,
//        symmetric, loop-free, duplicate-free graph
//   misfit_egr mis <graph.egr> <out.txt>
//        greedy ascending-id MIS, one 1-indexed id per line
//        (tools/compute_initial_mis.py)
//   misfit_egr batches <graph.egr> <outdir> <batch_size> <num_batches> [uniform|low-degree|preferential] [seed]
//        distinct existing edges in batch files <outdir>/<k>.mtx, "u v"
//        1-indexed with u < v (tools/generate_test_batches.py,
//        generate_low_degree_batches.py, generate_preferential_batches.py)
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/misfit.h"

static int usage() {
    std::cerr << "Usage: misfit_egr convert <input.mtx> <output.egr>\n"
                 "       misfit_egr check <graph.egr>\n"
                 "       misfit_egr mis <graph.egr> <out.txt>\n"
                 "       misfit_egr batches <graph.egr> <outdir> <batch_size> <num_batches>"
                 " [uniform|low-degree|preferential] [seed]" << std::endl;
    return 1;
}

// The graph with int64 offsets (narrow files are widened).
struct Csr {
    int32_t n = 0;
    int64_t m = 0;
    std::vector<int64_t> nindex;
    std::vector<int32_t> nlist;
};

static bool load(const char* path, Csr& g) {
    char err[512] = {0};
    misfit_egr e;
    if (misfit_egr_open(path, &e, err, sizeof(err)) != MISFIT_OK) {
        std::cerr << "ERROR: " << err << std::endl;
        return false;
    }
    g.n = e.n;
    g.m = e.m;
    g.nindex.resize((size_t)e.n + 1);
    for (int64_t i = 0; i <= e.n; i++)
        g.nindex[i] = e.wide ? ((const int64_t*)e.nindex)[i] : (int64_t)((const int32_t*)e.nindex)[i];
    g.nlist.assign(e.nlist, e.nlist + e.m);
    misfit_egr_close(&e);
    if (g.nindex[0] != 0 || g.nindex[g.n] != g.m) {
        std::cerr << "ERROR: " << path << ": offsets do not span the entry list" << std::endl;
        return false;
    }
    return true;
}

static int cmd_mis(const char* egr, const char* out) {
    Csr g;
    if (!load(egr, g)) return 1;
    std::cout << egr << ": " << g.n << " nodes, " << g.m << " directed edge entries" << std::endl;
    std::vector<uint8_t> mis((size_t)g.n);
    if (misfit_greedy_mis(g.n, g.nindex.data(), g.nlist.empty() ? nullptr : g.nlist.data(), mis.data()) !=
        MISFIT_OK) {
        std::cerr << "ERROR: greedy MIS failed" << std::endl;
        return 1;
    }
    FILE* f = fopen(out, "w");
    if (!f) {
        std::cerr << "ERROR: cannot write " << out << std::endl;
        return 1;
    }
    int64_t count = 0;
    for (int32_t v = 0; v < g.n; v++)
        if (mis[v]) {
            fprintf(f, "%d\n", v + 1);
            count++;
        }
    if (fclose(f) != 0) {
        std::cerr << "ERROR: cannot write " << out << std::endl;
        return 1;
    }
    printf("greedy MIS: %lld members (%.1f%% of nodes)\nwrote %s\n", (long long)count,
           g.n ? 100.0 * (double)count / g.n : 0.0, out);
    return 0;
}

static int cmd_batches(int argc, char** argv) {
    const char* egr = argv[2];
    std::string outdir = argv[3];
    long long bs = atoll(argv[4]), nb = atoll(argv[5]);
    std::string sampler = argc > 6 ? argv[6] : "uniform";
    unsigned long long seed = argc > 7 ? strtoull(argv[7], nullptr, 10) : 42ull;
    int kind = sampler == "uniform" ? 4 : sampler == "low-degree" ? 5 : sampler == "preferential" ? 6 : -1;
    if (kind < 0 || bs <= 0 || nb <= 0) return usage();
    Csr g;
    if (!load(egr, g)) return 1;
    std::cout << "Graph: " << g.n << " nodes, " << g.m << " directed CSR entries" << std::endl;
    const int64_t total = (int64_t)bs * nb;
    std::vector<int32_t> u((size_t)total), v((size_t)total);
    if (misfit_gen_updates(kind, g.n, g.nindex.data(), g.nlist.empty() ? nullptr : g.nlist.data(), total,
                           (uint64_t)seed, u.data(), v.data()) != MISFIT_OK) {
        std::cerr << "ERROR: could not draw " << total << " distinct edges from " << egr << std::endl;
        return 1;
    }
    mkdir(outdir.c_str(), 0755);
    for (long long k = 1; k <= nb; k++) {
        std::string path = outdir + "/" + std::to_string(k) + ".mtx";
        FILE* f = fopen(path.c_str(), "w");
        if (!f) {
            std::cerr << "ERROR: cannot write " << path << std::endl;
            return 1;
        }
        for (int64_t i = (k - 1) * bs; i < k * bs; i++) fprintf(f, "%d %d\n", u[i] + 1, v[i] + 1);
        if (fclose(f) != 0) {
            std::cerr << "ERROR: cannot write " << path << std::endl;
            return 1;
        }
        std::cout << "batch " << k << ": " << bs << " edges -> " << path << std::endl;
    }
    std::cout << "Wrote " << nb << " " << sampler << " batches (" << bs << " edges each) to " << outdir
              << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    std::string cmd = argv[1];
    char err[512] = {0};
    if (cmd == "convert" && argc == 4) {
        std::cout << "Converting " << argv[2] << " -> " << argv[3] << std::endl;
        int32_t n = 0;
        int64_t m = 0;
        if (misfit_mtx_to_egr(argv[2], argv[3], &n, &m, err, sizeof(err)) != MISFIT_OK) {
            std::cerr << "ERROR: " << err << std::endl;
            return 1;
        }
        std::cout << "Graph stats: " << n << " nodes, " << m << " edges." << std::endl;
        std::cout << "Conversion complete successfully!" << std::endl;
        return 0;
    }
    if (cmd == "check" && argc == 3) {
        misfit_egr_report r;
        if (misfit_egr_check(argv[2], &r, err, sizeof(err)) != MISFIT_OK) {
            std::cerr << "ERROR: " << err << std::endl;
            return 1;
        }
        std::cout << argv[2] << ": n=" << r.n << " m=" << r.m << " nindex=" << (r.wide ? "int64" : "int32")
                  << "\n  bad offsets:   " << r.bad_offsets << "\n  out of range:  " << r.out_of_range
                  << "\n  self-loops:    " << r.self_loops << "\n  duplicates:    " << r.duplicates
                  << "\n  unsorted rows: " << r.unsorted_rows << "\n  asymmetric:    "
                  << (r.asymmetric < 0 ? std::string("not checked") : std::to_string(r.asymmetric)) << std::endl;
        bool ok = !r.bad_offsets && !r.out_of_range && !r.self_loops && !r.duplicates && r.asymmetric == 0;
        std::cout << (ok ? "OK" : "NOT A VALID SIMPLE SYMMETRIC GRAPH") << std::endl;
        return ok ? 0 : 1;
    }
    if (cmd == "mis" && argc == 4) return cmd_mis(argv[2], argv[3]);
    if (cmd == "batches" && argc >= 6 && argc <= 8) return cmd_batches(argc, argv);
    return usage();
}
