// misfit_egr -- graph-file tool: .mtx -> .egr conversion and .egr checks.
//   misfit_egr convert <input.mtx> <output.egr>
//        what tools/efficient_converter.py does (wide form, file row order)
//   misfit_egr check <graph.egr>
//        structure report; exit status 1 when the file is not a valid,
//        symmetric, loop-free, duplicate-free graph
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/misfit.h"

static int usage() {
    std::cerr << "Usage: misfit_egr convert <input.mtx> <output.egr>\n"
                 "       misfit_egr check <graph.egr>" << std::endl;
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 4 && argc != 3) return usage();
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
    return usage();
}
