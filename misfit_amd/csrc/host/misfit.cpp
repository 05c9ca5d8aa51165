// misfit.cpp -- drop-in for the reference's synchronous binary
// (src/misfit/misfit.cu:732-963): batch i+1 starts only after batch i's
// device work has finished and been read back.
#include "driver.h"

int main(int argc, char** argv) {
    using namespace mfd;
    Args a = parse_args(argc, argv);
    std::cout << "=========================================START=====================================================" << std::endl;
    std::cout << "======================================= MISFIT (SYNC) =========================================" << std::endl;
    auto t_setup = clk::now();
    Setup S = setup(a);
    double setup_ms = ms_since(t_setup);
    std::cout << "Graph: " << a.egr << " nodes=" << S.n << " edges=" << S.m
              << " avg_degree=" << std::fixed << std::setprecision(2) << (S.n ? (double)S.m / S.n : 0.0) << std::endl;
    std::cout << "GPU: " << S.device_name << std::endl;
    std::cout << "Initial MIS cardinality: " << S.initial_card << std::endl;
    int nb = std::min<int>(a.num_batches, (int)S.batches.size());
    std::vector<double> lat;
    double time_cum = 0.0;
    for (int i = 0; i < nb; i++) {
        const Batch& b = S.batches[(size_t)i];
        std::cout << "\n" << std::string(60, '=') << std::endl;
        std::cout << "Processing batch " << (i + 1) << "/" << nb << " | Edges: " << b.ups.size() << std::endl;
        auto t0 = clk::now();
        misfit_batch_stats st;
        check(S.ctx, misfit_process_batch(S.ctx, b.ups.data(), (int64_t)b.ups.size(), &st), "batch");
        double ms = ms_since(t0);
        time_cum += ms;
        lat.push_back(ms);
        int64_t conf = 0, unc = 0, bad = 0;
        if (S.check_each) {
            check(S.ctx, misfit_check(S.ctx, &conf, &unc, &bad), "check");
            if (conf || bad) die("check failed after batch " + std::to_string(i + 1));
        }
        std::cout << "  Current MIS cardinality: [not computed per-batch]" << std::endl;
        print_batch_detail(i, nb, b, st, ms, S.check_each, conf, unc);
        sample_memory(S);
    }
    finish(S, a, lat, time_cum, setup_ms, false);
    return 0;
}
