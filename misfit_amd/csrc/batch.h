// batch.h -- parameters of one update batch on the device, shared by the
// batch kernels and the host code that launches them.
#pragma once
#include "dev.h"

// Rows longer than this get a per-batch index of their admissible
// neighbours instead of being scanned in full by the repair.  Admissible
// vertices are few, so every row uses the index.
#define MF_FA_T 0

// Rows of members that may be evicted are counted in chunks of this many
// entries, one wave per chunk, so hub rows do not serialise on one wave.
#define MF_KCH 64

// Guarded reads logged per wave and speculation round (see ReadLog).
#define MF_RLOG 128

enum BatchCounter {
    C_SEEDS = 0,     // distinct endpoints
    C_WPOT,          // distinct max(u,v) of insertions
    C_NE,            // new directed overlay entries
    C_PU,            // vertices that may become uncovered
    C_ADM,           // ... of which have a row (may be admitted)
    C_BIG,           // long rows with admissible neighbours
    C_FA_TOP,        // admissible-index entries
    C_ROUNDS,        // speculative rounds
    C_ACT_INS,       // insertions with both endpoints in the MIS
    C_EVICT,
    C_ADMIT_INS,
    C_ACT_DEL,       // deletions with exactly one endpoint in the MIS
    C_ADMIT_DEL,
    C_COMP_ADDS,
    C_COMP_ROUNDS,
    C_DELTA_OPS,     // effective overlay operations (delta_edge_count)
    C_POOL_OVF,
    C_FA_OVF,
    C_CAND,
    C_SPECS,         // edge evaluations (speculation work)
    C_NACT0, C_NACT1,
    C_CONF0, C_CONF1,
    C_UNDEC0, C_UNDEC1,
    C_START0, C_START1,  // first uncommitted edge, read by the round of that parity
    C_KTASK,         // row chunks of members that may be evicted
    C_GUARD,         // a loop bound was hit or a list overflowed (engine bug): results invalid
    C_HALT,          // the tail stopped with many edges left: the host queues more rounds
    C_T0, C_TP, C_T1,  // device wall clock: batch start, first round, tail end
    C_TR, C_TS,        // ... the tail's last round done, its rule counters done
    C_NCOUNTERS
};

// Each counter sits on a cache line of its own (16 x 8 bytes): waves that
// append to one list must not stall the reads of another list's length.
#define MF_CS 16

struct BP {
    DevGraph g;
    const int32_t* bu;
    const int32_t* bv;
    const uint8_t* bins;
    // the batch as handed over (pinned host memory, read by the first
    // kernel into bu/bv/bins), or null when bu/bv/bins already hold it
    const int32_t* hu;
    const int32_t* hv;
    const uint8_t* hins;
    int64_t U;
    uint32_t b;        // batch epoch, >= 1
    int completion;
    int64_t window;    // max edges speculated per round
    int64_t fin_window;// max edges per round of the one-workgroup tail (0: the default)
    int64_t n_ins;     // insertion updates in the batch (bounds evictions per round)
    int may_halt;      // the tail may stop with edges left (nothing of this graph is queued behind)
    int pu_list;       // the MIS was maximal before the batch: the PU scan tests only tagged vertices
    // pair table of this batch
    u64* bt_key;
    int32_t* bt_head;
    int32_t* bt_nxt;
    int32_t* bt_slot;
    u64 bt_mask;
    // pre-batch adjacency of update i's pair, searched beside the claim:
    // positions of hi in row lo and of lo in row hi (-1: not in the base),
    // bit 0 = live overlay entry, bit 1 = base entry deleted; P2 adds
    // bit 2 = adjacent after the batch
    int64_t* bt_plo;
    int64_t* bt_phi;
    uint8_t* bt_fl;
    // new overlay entries
    int32_t* ne_x;
    int32_t* ne_y;
    int32_t* ne_seq;
    int32_t* ne_nxt;
    // vertex lists
    int32_t* seeds;
    int32_t* wpot;
    u64* ktask;        // w<<32 | chunk
    int64_t ktask_cap;
    int32_t* pu;
    int64_t pu_cap;
    int32_t* big;
    int64_t big_cap;
    int32_t* fa_z;
    int32_t* fa_pos;
    int64_t fa_cap;
    // repair
    int32_t* active;
    uint8_t* kind;
    int32_t* res_w;
    int32_t* res_x;
    uint4* rlog;       // guarded reads of a round, MF_RLOG slots per wave of the grid (slot 0: header)
    int4* cm;          // per wave of the round grid: its first edge's kind, evicted
                       // and admitted vertex and index, left by the speculation for the commit
    int4* stage;       // per wave of the round grid: the endpoints, batch epoch and
                       // insertion flag (epoch << 1 | ins) and index of the edge it speculates
                       // next round, left by the commit that fixes where that round starts
                       // (index -1: none)
    int64_t nstage;    // entries of stage (waves of the largest round grid)
    uint32_t* relog;   // per edge: = round when its read log overflowed
    // completion
    int32_t* cand;
    uint8_t* cdec;
    u64* ctr;          // counter i at ctr[MF_CS * i]
    uint32_t* round;   // global speculative-round counter (device-resident)
    u64* h_out;        // pinned host copy of the counters + pool top, written
                       // by the batch's last kernel (C_NCOUNTERS + 1 words)
};

// The serving kernel's work queue, in pinned host memory (coherent): the
// host writes an item, then raises `tail`; a workgroup of the kernel claims
// the next item with a compare-and-swap on a head counter in device memory,
// runs the batch's speculative rounds, rule counters and completion (and
// its prologue too for a `whole` item; otherwise the prologue's kernels ran
// on the graph's own stream before the item was published), publishes its
// counters to P.h_out and then writes the item's index plus one to
// P.h_out[MF_DONE_WORD].
#define MF_SERVE_CAP 4096
struct ServeItem {
    BP P;
    u64 whole;      // 1: the workgroup runs the prologue too; 0: it ran on the graph's stream
    u64 pad[4];
};
struct ServeQueue {
    u64 tail;       // items published
    u64 stop;       // nonzero: every workgroup leaves at its next claim
    u64 pad[14];
    ServeItem items[MF_SERVE_CAP];
};
// Word of P.h_out that reports a finished batch of the serving kernel.
#define MF_DONE_WORD (C_NCOUNTERS + 1)
#define MF_HOUT_WORDS (C_NCOUNTERS + 2)

// host-side launchers (k_batch.hip)
struct LaunchCfg {
    int blocks;        // grid cap for grid-stride phases (8 workgroups per CU)
    hipStream_t stream;
    int round_per_cu = 0;  // speculation/validation workgroups per CU (0: by occupancy)
    // Per-batch sizing (U < 0: every phase gets the whole GPU).  Phases get
    // as many workgroups as their work needs, so a small batch leaves the
    // GPU to other graphs' batches: a thread per update, a wave per update,
    // a thread per four vertices (or per 32 when only tagged ones are
    // tested); the rounds get round_grid workgroups.
    int64_t U = -1;
    int64_t n = 0;
    int round_grid = 0;
    int round_threads = 256;  // workgroup size of the speculation, validation and commit kernels: 128, 256 or 512
};
// Waves per workgroup of the round kernels (a round's window is round_grid times this).
inline int mf_round_wpg(const LaunchCfg& L) { return L.round_threads / MF_WAVE; }
void mf_launch_batch_prologue(const BP& P, const LaunchCfg& L);
int mf_round_grid(const LaunchCfg& L);  // workgroups of the speculation/validation kernels
void mf_launch_round(const BP& P, const LaunchCfg& L, int par, bool first = false);  // first: stamps C_TP
void mf_launch_finish(const BP& P, const LaunchCfg& L, int par);
void mf_launch_completion_collect(const BP& P, const LaunchCfg& L);
void mf_launch_stats(const BP& P, const LaunchCfg& L);
void mf_launch_completion_round(const BP& P, const LaunchCfg& L, int par);
// The serving kernel: `grid` workgroups of `threads` threads (256, 512 or
// 1024), each taking whole batches from the queue (q and head as the device
// addresses them) until the host sets stop or no item comes for idle_ticks
// of the device clock.
void mf_launch_serve(ServeQueue* q, u64* head, int grid, int threads, u64 idle_ticks, hipStream_t st);
void mf_launch_publish(const BP& P, const LaunchCfg& L);
