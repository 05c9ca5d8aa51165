// graph.h -- host-side handles of one device-resident graph and the
// whole-graph launchers of k_graph.hip.
#pragma once
#include "batch.h"

struct GraphState {
    DevGraph g;
    hipStream_t stream;
    int blocks;              // grid for grid-stride kernels
    uint32_t* d_cnt;         // n: recount scratch
    uint32_t* d_bmark;       // n: byte-measurement marks
    int64_t* d_deg;          // n: compaction degrees
    int64_t* d_scan_part;    // scan block sums
    void* d_walk_tasks;      // edge-walk split tasks
    u64* d_walk_ntasks;
    int64_t walk_task_cap;
};

// Frees device or pinned host memory.  hipFree and hipHostFree wait for
// every kernel on the device to end, on any stream, the long-running serving
// kernel included; while a serving kernel may run, frees are therefore
// deferred until it has ended (the memory is only released later, never
// used again).
void mf_free_device(void* p);
void mf_free_host(void* p);
void mf_serving_begin();  // before a serving kernel's launch
void mf_serving_end();    // after its end was seen: deferred frees run once none is left

// Split tasks of the edge walk (rows cut into pieces of 4,096 entries) a base
// of n vertices and m entries may need.
inline int64_t mf_walk_task_need(int64_t n, int64_t m) { return m / 4096 + n / 64 + 1024; }

hipError_t mf_set_mis(GraphState& S, const uint8_t* d_mis);
// acc[0] conflicts, acc[1] uncovered, acc[2] count errors, acc[3] walk overflow (results invalid)
hipError_t mf_get_mis(GraphState& S, uint8_t* d_mis);
hipError_t mf_flags(GraphState& S, const int32_t* d_u, const int32_t* d_v, int64_t count, uint8_t* d_ins);
hipError_t mf_check(GraphState& S, u64* d_acc);
hipError_t mf_compact_plan(GraphState& S, int64_t* d_newidx);
// Deletion marks of 64 consecutive entries of an array of rows (old entries
// in row order, or the old sorted copy): bit i set = entry 64w+i is a
// deleted base edge; pre = deleted entries before entry 64w in the whole
// array.  pre - (deleted entries in rows before x) = deleted entries of row
// x before the word, so one 16-byte load answers "is entry p deleted" and
// "how many deletions precede it in its row".
struct __align__(16) TombWord {
    u64 mask;
    u64 pre;
};
// words covering m entries (one past the end, so the end of the last row has a word)
__host__ __device__ inline int64_t mf_tomb_words(int64_t m) { return m / 64 + 1; }
// entries per counting chunk of the position-order marks (a wave's task; a multiple of 64)
#define MF_TCHUNK 8192
__host__ __device__ inline int64_t mf_tomb_chunks(int64_t m) { return (mf_tomb_words(m) * 64 + MF_TCHUNK - 1) / MF_TCHUNK; }
// Merge-path chunks of an array of rows: chunk t starts where rows passed
// plus entries passed reach t * MF_MCHUNK (rounded down to a multiple of 64
// entries), so a chunk holds at most about MF_MCHUNK entries and crosses at
// most about MF_MCHUNK rows, however sparse the rows are.
#define MF_MCHUNK 8192
__host__ __device__ inline int64_t mf_path_chunks(int64_t rows, int64_t entries) {
    return (rows + entries + MF_MCHUNK - 1) / MF_MCHUNK;
}

// What a compaction hands from the copy to the merge of the sorted row copy.
struct MergeLists {
    const int64_t* doff = nullptr;   // n+1: deleted base entries in rows before x (exclusive scan)
    int32_t* tval = nullptr;         // live overlay entries, insertion order
    int32_t* tidx = nullptr;         // ... their index in that order
    const int64_t* toff = nullptr;   // n+1: offsets of the rows' tails
    const int32_t* trows = nullptr;  // rows with a tail
    const u64* ntrows = nullptr;
    TombWord* tw = nullptr;          // deletion marks of the old rows (position order)
    TombWord* tws = nullptr;         // ... of the old sorted copy (null: rows were ascending, use tw)
    int64_t* ccnt = nullptr;         // per counting chunk: deletions in it
    int64_t* cpre = nullptr;         // ... exclusive scan (chunks + 1)
    int64_t* cpart = nullptr;        // scan scratch
    int64_t* cst = nullptr;          // merge-path chunk starts of the old rows (ncst + 1 entries)
    int64_t ncst = 0;
    // tail sort: sorted tail values and their insertion indices, rows whose
    // tail is long, a counter, and the library sort's temporary storage
    int32_t* tsv = nullptr;
    int32_t* tsi = nullptr;
    int32_t* big = nullptr;
    u64* nbig = nullptr;
    void** tmp = nullptr;
    size_t* tmp_bytes = nullptr;
};
// Deletion marks of the old rows (L.tw), then the kept base entries of every
// row into d_out at their compacted positions, then the live overlay
// entries after them (collected into L.tval/L.tidx).  flags |= 1 when a
// compacted row may not be ascending (the old rows were not, or a row's
// tail does not continue it in ascending order).
hipError_t mf_compact_copy(GraphState& S, const int64_t* d_newidx, int32_t* d_out, const MergeLists& L, int* d_flags);
hipError_t mf_batch_bytes(GraphState& S, const int32_t* d_seeds, const u64* d_nseeds, uint32_t t, u64* d_acc);
hipError_t mf_widen(const int32_t* d_in, int64_t* d_out, int64_t count, hipStream_t st);
hipError_t mf_validate_csr(GraphState& S, int* d_flags);
hipError_t mf_validate_entries(GraphState& S, int* d_flags);
// sorted copy of every row with positions (rows not ascending); flags |= 16 on a repeated neighbour,
// |= 32 when the copy does not match the rows
hipError_t mf_build_row_index(GraphState& S, int32_t* sval, int32_t* spos, int* d_flags);
// the overlay's deletions and tails before each row (doff, toff) and the new
// row starts (nidx), n+1 each, in one scan; the rows with a tail listed in
// trows in row order (*ntrows of them); part: scan scratch of 3 * ceil(n / 2048) entries
hipError_t mf_compact_plan_counts(GraphState& S, int64_t* nidx, int64_t* doff, int64_t* toff, int32_t* trows,
                                  u64* ntrows, int64_t* part);
// sorted copy of the compacted rows nidx/nnl by merging the old graph's (S) sorted order with the
// tails compaction collected, queued on S's stream without waiting; with `check`, the result is
// checked entry by entry (the flags of mf_build_row_index, ORed into *d_flags)
hipError_t mf_merge_row_index(GraphState& S, const MergeLists& L, int64_t ntail, const int64_t* nidx,
                              const int32_t* nnl, int32_t* nsval, int32_t* nspos, int* d_flags, bool check);
void mf_scan_excl(const int64_t* in, int64_t count, int64_t* out, int64_t* part, hipStream_t st);
void mf_fill_u32(uint32_t* p, uint32_t v, int64_t count, hipStream_t st);
void mf_fill_i32(int32_t* p, int32_t v, int64_t count, hipStream_t st);
void mf_fill_u64(u64* p, u64 v, int64_t count, hipStream_t st);
void mf_fill_i64(int64_t* p, int64_t v, int64_t count, hipStream_t st);
