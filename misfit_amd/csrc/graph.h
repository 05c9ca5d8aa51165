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
    int32_t* d_tomb_list;    // n: rows with deletions
    int64_t* d_scan_part;    // scan block sums
    void* d_walk_tasks;      // edge-walk split tasks
    u64* d_walk_ntasks;
    int64_t walk_task_cap;
};

hipError_t mf_set_mis(GraphState& S, const uint8_t* d_mis);
hipError_t mf_get_mis(GraphState& S, uint8_t* d_mis);
hipError_t mf_flags(GraphState& S, const int32_t* d_u, const int32_t* d_v, int64_t count, uint8_t* d_ins);
hipError_t mf_check(GraphState& S, u64* d_acc);
hipError_t mf_compact_plan(GraphState& S, int64_t* d_newidx);
hipError_t mf_compact_copy(GraphState& S, const int64_t* d_newidx, int32_t* d_out);
hipError_t mf_batch_bytes(GraphState& S, const int32_t* d_seeds, const u64* d_nseeds, uint32_t t, u64* d_acc);
hipError_t mf_widen(const int32_t* d_in, int64_t* d_out, int64_t count, hipStream_t st);
hipError_t mf_validate_csr(GraphState& S, int* d_flags);
hipError_t mf_validate_entries(GraphState& S, int* d_flags);
// sorted copy of every row with positions (rows not ascending); flags |= 16 on a repeated neighbour
hipError_t mf_build_row_index(GraphState& S, int32_t* sval, int32_t* spos, int* d_flags);
void mf_scan_excl(const int64_t* in, int64_t count, int64_t* out, int64_t* part, hipStream_t st);
void mf_fill_u32(uint32_t* p, uint32_t v, int64_t count, hipStream_t st);
void mf_fill_i32(int32_t* p, int32_t v, int64_t count, hipStream_t st);
void mf_fill_u64(u64* p, u64 v, int64_t count, hipStream_t st);
void mf_fill_i64(int64_t* p, int64_t v, int64_t count, hipStream_t st);
