// ref_driver.cpp -- exposes the reference's own host classes through a C ABI
// so the tests can check the oracle's overlay and compaction against them.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference's unmodified headers (core/FlashGraph.h,
// core/flash_compactor.h), which are included from where they lie in the
// reference checkout (REF_CORE on the compiler line).  Nothing in the product
// links or loads this.
#include <cstdint>
#include <cstring>
#include <vector>
#include "FlashGraph.h"
#include "flash_compactor.h"

extern "C" {

__attribute__((visibility("default"))) void* ref_open(const char* path) {
    HybridGraph* g = new HybridGraph();
    g->open(path);
    return g;
}
__attribute__((visibility("default"))) void ref_close(void* h) { delete (HybridGraph*)h; }
__attribute__((visibility("default"))) int32_t ref_nodes(void* h) { return ((HybridGraph*)h)->num_nodes; }
__attribute__((visibility("default"))) int64_t ref_base_edges(void* h) { return (int64_t)((HybridGraph*)h)->base.edges; }
__attribute__((visibility("default"))) int64_t ref_base_nindex(void* h, int32_t v) { return ((HybridGraph*)h)->base.nindex_at(v); }
__attribute__((visibility("default"))) int ref_is_adjacent(void* h, int32_t u, int32_t v) {
    return ((HybridGraph*)h)->isAdjacent(u, v) ? 1 : 0;
}
__attribute__((visibility("default"))) void ref_insert(void* h, int32_t u, int32_t v) { ((HybridGraph*)h)->insertEdge(u, v); }
__attribute__((visibility("default"))) void ref_remove(void* h, int32_t u, int32_t v) { ((HybridGraph*)h)->removeEdge(u, v); }
__attribute__((visibility("default"))) int64_t ref_degree(void* h, int32_t v) { return ((HybridGraph*)h)->degree(v); }
__attribute__((visibility("default"))) int64_t ref_delta_ops(void* h) { return ((HybridGraph*)h)->delta.delta_edge_count; }
__attribute__((visibility("default"))) int64_t ref_neighbors(void* h, int32_t v, int32_t* out, int64_t cap) {
    std::vector<int> nb;
    ((HybridGraph*)h)->neighbors_into(v, nb);
    int64_t n = (int64_t)nb.size();
    if (out) std::memcpy(out, nb.data(), (size_t)(n < cap ? n : cap) * sizeof(int32_t));
    return n;
}
// A batch applied as the reference's synchronous engine applies it to its
// overlay (src/misfit/misfit.cu:708-714): insert when flagged as an
// insertion and not adjacent, delete when flagged as a deletion and adjacent.
// check = 0 skips the adjacency test (a linear scan of the base row in the
// reference), which gives the same overlay when every update changes the
// graph: deletions of present edges and insertions of absent ones, each pair
// once.
__attribute__((visibility("default"))) void ref_apply(void* h, const int32_t* u, const int32_t* v, const uint8_t* ins,
                                                      int64_t count, int check) {
    HybridGraph* g = (HybridGraph*)h;
    for (int64_t i = 0; i < count; i++) {
        if (ins[i]) {
            if (!check || !g->isAdjacent(u[i], v[i])) g->insertEdge(u[i], v[i]);
        } else if (!check || g->isAdjacent(u[i], v[i])) {
            g->removeEdge(u[i], v[i]);
        }
    }
}
// Compacts base + overlay into out_path with the reference's writer.
__attribute__((visibility("default"))) void ref_compact(void* h, const char* out_path) {
    HybridGraph* g = (HybridGraph*)h;
    compact_flash_graph(g->base, g->delta, out_path);
}

}  // extern "C"
