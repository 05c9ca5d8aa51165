// k_graph.hip -- whole-graph kernels: load, MIS upload with the
// MIS-neighbour counts, adjacency flags, the full independence/maximality
// check, compaction of the overlay into a new base, and the per-batch
// algorithmic-byte measurement.
//
// Whole-graph passes walk the base CSR edge-parallel: a wave owns 64
// consecutive vertices, whose rows are contiguous in nlist, and its lanes read
// that range 64 entries at a time (coalesced); each lane finds the row of its
// entry by a 6-step search over the 64 row offsets held in registers.  Ranges
// longer than MF_SPLIT are cut into tasks that other waves pick up, so hub
// rows do not serialise a pass.  Overlay runs are walked per vertex.
#include <algorithm>
#include <vector>

#include "graph.h"

#define MF_SPLIT 4096

// Entries per wave task of the entry-balanced passes (entry_walk).
#define MF_ECHUNK 8192

// Visits every entry p of rows [r0, r1) once, balanced by entries: the
// entries are cut into tasks of MF_ECHUNK, one wave per task; the wave finds
// the row of its first entry by binary search, then walks windows of 64
// consecutive rows (a lane holds one row's start and rv(row), a per-row
// value) and finds each entry's row among them by a 6-step shuffle search.
// Hub rows are spread over many waves, short rows cost no search each.
// visit(p, row, row start, rv(row)) runs on the lanes that hold an entry;
// it must not use cross-lane operations.
template <class RowVal, class Visit>
__device__ __forceinline__ void entry_walk(const int64_t* nindex, int32_t r0, int32_t r1, RowVal rv, Visit visit) {
    GC c = grid_coords();
    const int64_t pb = nindex[r0], pe = nindex[r1];
    const int64_t ntask = (pe - pb + MF_ECHUNK - 1) / MF_ECHUNK;
    for (int64_t t = c.wave; t < ntask; t += c.nwaves) {
        int64_t p = pb + t * MF_ECHUNK;
        const int64_t tend = p + MF_ECHUNK < pe ? p + MF_ECHUNK : pe;
        int64_t lo = r0, hi = r1 - 1;  // largest row x with nindex[x] <= p
        while (lo < hi) {
            int64_t mid = (lo + hi + 1) >> 1;
            if (nindex[mid] <= p) lo = mid; else hi = mid - 1;
        }
        int64_t x = lo;
        while (p < tend && x < r1) {  // x grows every window: the walk ends
            const int64_t myv = x + c.lane;
            const int64_t off = myv < r1 ? nindex[myv] : INT64_MAX;
            const int64_t val = myv < r1 ? rv(myv) : 0;
            const int64_t xw = x + MF_WAVE < r1 ? x + MF_WAVE : r1;
            const int64_t wend = nindex[xw];
            const int64_t lim = wend < tend ? wend : tend;
            for (int64_t p0 = p; p0 < lim; p0 += MF_WAVE) {
                const int64_t pp = p0 + c.lane;
                const int64_t q = pp < lim ? pp : lim - 1;
                int l = 0, h = MF_WAVE - 1;
#pragma unroll
                for (int s = 0; s < 6; s++) {
                    int mid = (l + h + 1) >> 1;
                    int64_t om = __shfl(off, mid, MF_WAVE);
                    if (om <= q) l = mid; else h = mid - 1;
                }
                const int64_t ostart = __shfl(off, l, MF_WAVE);
                const int64_t oval = __shfl(val, l, MF_WAVE);
                if (pp < lim) visit(pp, x + l, ostart, oval);
            }
            if (lim > p) p = lim;
            x = xw;
        }
    }
}

// Largest row x in [r0, r1) with nindex[x] <= p (nindex[r0] <= p), by a
// 64-ary search in which every lane probes one position per step: 4
// dependent loads for 2^24 rows instead of 24.  Wave-uniform result; all
// lanes of the wave must call it together.
__device__ __forceinline__ int64_t row_of(const int64_t* nindex, int64_t r0, int64_t r1, int64_t p, int lane) {
    int64_t lo = r0, hi = r1 - 1;  // the answer lies in [lo, hi]
    while (lo < hi) {
        const int64_t span = hi - lo;
        const int64_t probe = lo + ((int64_t)(lane + 1) * span + 63) / 64;  // ascending, lane 63 at hi
        const u64 le = __ballot(nindex[probe] <= p);
        const int k = __popcll(le);  // probes 0 .. k-1 are <= p
        const int64_t plo = __shfl(probe, k > 0 ? k - 1 : 0, MF_WAVE);
        const int64_t phi = __shfl(probe, k < 64 ? k : 63, MF_WAVE);
        if (k > 0) lo = plo;
        if (k < 64) hi = phi - 1;
    }
    return lo;
}

// Value of v on lane l (l wave-uniform): scalar reads, no shuffle.
__device__ __forceinline__ int64_t read_lane(int64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((u64)v >> 32), l);
    return (int64_t)(((u64)hi << 32) | lo);
}
__device__ __forceinline__ int32_t read_lane(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Inclusive prefix maximum over the lanes of the wave, in lane order, of
// values >= 0: within each row of 16 lanes by shifts of 1, 2, 4 and 8, then
// across rows by the broadcasts of lanes 15 and 31 (DPP, no LDS traffic).
__device__ __forceinline__ int wave_prefix_max(int v) {
    int t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v = t > v ? t : v;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v = t > v ? t : v;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v = t > v ? t : v;
    t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v = t > v ? t : v;
    t = __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1 and 3
    v = t > v ? t : v;
    t = __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2 and 3
    v = t > v ? t : v;
    return v;
}

// A row of the current window as the walk's table holds it in LDS: its
// first entry and the per-row values its entries use.
template <class Pay>
struct RowRec {
    int64_t start;
    Pay pay;
};

// One step of an entry walk: MF_EW_GROUPS groups of 64 consecutive entries
// from p0, group j's lane l holding entry p0 + 64 j + l.  Everything but
// the row lanes is wave-uniform; rows are found through the window's table.
#define MF_EW_GROUPS 4
template <class Pay>
struct WalkStep {
    int64_t p0, lim;        // the step's first entry; entries from lim on are not this step's
    int64_t x;              // the window's first row
    int64_t chunk;          // the merge-path chunk of these entries
    RowRec<Pay>* tab;       // the window's rows (LDS)
    int rl[MF_EW_GROUPS];   // per lane: the window row (0..63) of its entry in each group
    __device__ __forceinline__ int64_t pp(int j) const { return p0 + (int64_t)(MF_WAVE * j + lane_id()); }
    __device__ __forceinline__ bool valid(int j) const { return pp(j) < lim; }
    // pp clamped into the step: always an entry index, for unconditional loads
    __device__ __forceinline__ int64_t q(int j) const {
        const int64_t e = pp(j);
        return e < lim ? e : lim - 1;
    }
    __device__ __forceinline__ int64_t start(int j) const { return tab[rl[j]].start; }
    __device__ __forceinline__ const Pay& pay(int j) const { return tab[rl[j]].pay; }
};

// Visits every entry of rows [r0, r1) once, balanced by merge-path chunks
// (cst, nch), for kernels bound by memory.  A wave walks its chunk through
// windows of 64 consecutive rows: lane l loads row x + l's start and
// rowpay(x + l), and the window's rows go to a table in LDS (two per wave,
// used in turn by the windows that have entries, so the step still
// waiting for its stores keeps its own).  Each step takes MF_EW_GROUPS
// groups of 64 entries.  The step's first row is the count of window rows
// starting at or before it; when the step spans several rows, every row
// that starts inside it puts its window index at its start in a 256-slot
// LDS array (max, so the last of several empty rows wins), and a prefix
// maximum over the slots gives each entry its row.  Steps are
// software-pipelined in three phases,
//     load1(next step)  ->  store(this step)  ->  load2(next step),
// where load2 may depend on load1's values and store on both.  On this
// GPU one counter tracks loads and stores together and in order, so a
// wave waiting for a load also waits for every older store: issuing the
// next step's loads before this step's stores keeps the store latency
// out of every wait.  The next window's row data is also loaded while the
// current window is walked.  Phases run on every lane (so `store` may use
// ballots); lanes past the step's end hold no entry (valid() is false) and
// an arbitrary window row.  Workgroups of MF_BLOCK threads.
template <class Pay, class RowPay, class Load1, class Load2, class Store>
__device__ __forceinline__ void entry_walk4(const int64_t* nindex, int32_t r0, int32_t r1, const int64_t* cst,
                                            int64_t nch, RowPay rowpay, Load1 load1, Load2 load2, Store store) {
    constexpr int STEP = MF_WAVE * MF_EW_GROUPS;
    __shared__ RowRec<Pay> tabs[MF_BLOCK / MF_WAVE][2][MF_WAVE];
    __shared__ __align__(16) int marks[MF_BLOCK / MF_WAVE][STEP];
    static_assert(STEP == 4 * MF_WAVE, "one int4 of marks per lane");
    GC c = grid_coords();
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / MF_WAVE));
    const int64_t wave0 = __builtin_amdgcn_readfirstlane((int)c.wave);
    int* mark = marks[wv];
    using D = decltype(load1(WalkStep<Pay>{}));
    WalkStep<Pay> cur;
    D dcur{};
    bool have = false;  // wave-uniform: cur holds a step whose stores are due
    int slot = 1;       // the table of the latest window with entries
    for (int64_t t = wave0; t < nch; t += c.nwaves) {
        int64_t p = cst[t];
        const int64_t tend = cst[t + 1];
        if (p >= tend) continue;  // a chunk of empty rows only
        int64_t x = read_lane(row_of(nindex, r0, r1, p, c.lane), 0);
        // window loads at clamped indices, selected after
        auto win = [&](int64_t x0, int64_t& off, Pay& pay, int64_t& wend) {
            const int64_t myv = x0 + c.lane;
            const int64_t lo_ = nindex[myv < r1 ? myv : r1];
            const Pay py = rowpay(myv < r1 ? myv : r1 - 1);
            const int64_t xw = x0 + MF_WAVE < r1 ? x0 + MF_WAVE : r1;
            wend = nindex[xw];
            off = myv < r1 ? lo_ : INT64_MAX;
            pay = myv < r1 ? py : Pay{};
        };
        int64_t off, wend;
        Pay pay;
        win(x, off, pay, wend);
        while (p < tend && x < r1) {
            const int64_t xw = x + MF_WAVE < r1 ? x + MF_WAVE : r1;
            int64_t noff = INT64_MAX, nwend = 0;
            Pay npay{};
            const int64_t lim = wend < tend ? wend : tend;
            if (lim < tend && xw < r1) win(xw, noff, npay, nwend);  // wave-uniform: the chunk goes on
            if (lim > p) {  // wave-uniform: the window has entries
                slot ^= 1;
                tabs[wv][slot][c.lane] = RowRec<Pay>{off, pay};
            }
            RowRec<Pay>* tab = tabs[wv][slot];
            for (int64_t p0 = p; p0 < lim; p0 += STEP) {
                WalkStep<Pay> nx;
                nx.p0 = p0;
                nx.lim = lim;
                nx.x = x;
                nx.chunk = t;
                nx.tab = tab;
                // the step's first row: the last window row starting at or before p0
                const int lf = __popcll(__ballot(off <= p0)) - 1;
                const int64_t rend = lf < MF_WAVE - 1 ? read_lane(off, lf + 1) : wend;
                if (p0 + STEP <= rend) {  // the whole step lies in row lf
#pragma unroll
                    for (int j = 0; j < MF_EW_GROUPS; j++) nx.rl[j] = lf;
                } else {
                    reinterpret_cast<int4*>(mark)[c.lane] = make_int4(0, 0, 0, 0);
                    const int64_t rel = off - p0;
                    if (rel >= 1 && rel < STEP) atomicMax(&mark[rel], c.lane);
                    int carry = lf;
#pragma unroll
                    for (int j = 0; j < MF_EW_GROUPS; j++) {
                        int s = wave_prefix_max(mark[MF_WAVE * j + c.lane]);
                        s = s > carry ? s : carry;
                        nx.rl[j] = s;
                        carry = __builtin_amdgcn_readlane(s, MF_WAVE - 1);
                    }
                }
                D dn = load1(nx);
                if (have) store(cur, dcur);
                load2(nx, dn);
                cur = nx;
                dcur = dn;
                have = true;
            }
            if (lim > p) p = lim;
            x = xw;
            off = noff;
            pay = npay;
            wend = nwend;
        }
    }
    if (have) store(cur, dcur);
}

// Merge-path chunk starts of rows [r0, r1) (see MF_MCHUNK): cst[t] for
// t = 0 .. nch, cst[nch] = the rows' end.  Chunk t starts on the largest
// row x whose (x - r0) + (nindex[x] - nindex[r0]) <= t * MF_MCHUNK, at the
// entry that brings the sum to t * MF_MCHUNK, rounded down to a multiple
// of 64 entries.
__global__ void k_merge_path(const int64_t* nindex, int32_t r0, int32_t r1, int64_t nch, int64_t* cst) {
    const int64_t pb = nindex[r0], pe = nindex[r1];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= nch; t += (int64_t)gridDim.x * blockDim.x) {
        if (t == nch) {
            cst[t] = pe;
            continue;
        }
        const int64_t d = t * MF_MCHUNK;
        int64_t lo = r0, hi = r1 - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if ((mid - r0) + (nindex[mid] - pb) <= d) lo = mid;
            else hi = mid - 1;
        }
        int64_t j = d - (lo - r0);
        if (j > pe - pb) j = pe - pb;
        cst[t] = pb + (j & ~(int64_t)63);
    }
}

static void merge_path(const int64_t* nindex, int32_t r0, int32_t r1, int64_t nch, int64_t* cst, hipStream_t st) {
    const int64_t nb = (nch + 1 + 255) / 256;
    k_merge_path<<<(unsigned)std::min<int64_t>(nb, 4096), 256, 0, st>>>(nindex, r0, r1, nch, cst);
}

struct PayNone {};
struct PayI64 {
    int64_t v;
};

// ---------------------------------------------------------------- basics
__global__ void k_widen32(const int32_t* in, int64_t* out, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

template <typename T>
__global__ void k_fill(T* p, T v, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// flags[0] |= 1: some row not strictly ascending; |= 2: self-loop;
// |= 4: entry out of range; |= 8: nindex not monotone.  Offsets are checked
// per vertex, entries per entry (the row of entry p by binary search), so
// long rows do not serialise the pass.
__global__ void k_validate_offsets(const int64_t* nindex, int32_t n, int64_t m, int* flags) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = nindex[x], e = nindex[x + 1];
        if (b > e || b < 0 || e > m) atomicOr(flags, 8);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (nindex[0] != 0 || nindex[n] != m)) atomicOr(flags, 8);
}

__global__ __launch_bounds__(256) void k_validate_entries(const int64_t* nindex, const int32_t* nlist, int32_t n,
                                                          int64_t m, int* flags) {
    int f = 0;
    (void)m;
    entry_walk(
        nindex, 0, n, [](int64_t) { return (int64_t)0; },
        [&](int64_t p, int64_t x, int64_t start, int64_t) {
            int32_t y = nlist[p];
            if (y < 0 || y >= n) f |= 4;
            if (y == (int32_t)x) f |= 2;
            if (p > start && nlist[p - 1] >= y) f |= 1;
        });
    if (f) atomicOr(flags, f);
}

__global__ void k_set_mis(uint32_t* mc, const uint8_t* mis, int32_t n) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        mc[x] = mis[x] ? MF_MISBIT : 0u;
}

__global__ void k_get_mis(const uint32_t* mc, uint8_t* mis, int32_t n) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        mis[x] = (mc[x] & MF_MISBIT) ? 1 : 0;
}

__global__ void k_flags(DevGraph g, const int32_t* u, const int32_t* v, int64_t count, uint8_t* ins) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        ins[i] = adjacent(g, u[i], v[i]) ? 0 : 1;
}

// ---------------------------------------------------------------- edge walk
// Visitor interface: visit(x, y) for every live entry y of row(x), y != x.
struct PassCountMis {      // cnt[x] += [y in MIS]
    const uint32_t* mc;
    uint32_t* cnt;
    u64* acc;              // acc[0] += conflicting entries (x<y, both members)
    __device__ void visit(int32_t x, int32_t y) const {
        uint32_t my = mc[y];
        if (my & MF_MISBIT) {
            atomicAdd(&cnt[x], 1u);
            if (acc && x < y && (mc[x] & MF_MISBIT)) atomicAdd(&acc[0], 1ull);
        }
    }
};

struct PassMarkV2 {        // mark[y] = t for y adjacent to a V1 vertex
    uint32_t* mark;
    uint32_t v1tag, v2tag;
    __device__ void visit(int32_t x, int32_t y) const {
        if (mark[x] != v1tag) return;
        if (mark[y] != v1tag) atomicMax(&mark[y], v2tag);
    }
};

struct EdgeTask {
    int32_t v0;
    int64_t p0;
};

template <class Vis>
__device__ void walk_range(const DevGraph& g, int32_t v0, int32_t v1, int64_t p_begin, int64_t p_end,
                           const Vis& vis) {
    // offsets of the (up to) 64 rows [v0, v1) in registers, one per lane
    int lane = lane_id();
    int32_t myv = v0 + lane;
    int64_t off = (myv < v1) ? g.nindex[myv] : INT64_MAX;
    for (int64_t p0 = p_begin; p0 < p_end; p0 += MF_WAVE) {
        int64_t p = p0 + lane;
        int64_t pp = p < p_end ? p : p_end - 1;
        int lo = 0, hi = MF_WAVE - 1;
#pragma unroll
        for (int s = 0; s < 6; s++) {
            int mid = (lo + hi + 1) >> 1;
            int64_t om = __shfl(off, mid, MF_WAVE);
            if (om <= pp) lo = mid; else hi = mid - 1;
        }
        if (p < p_end) {
            int32_t x = v0 + lo;
            int32_t y = g.nlist[p];
            if (y >= 0 && y != x) vis.visit(x, y);
        }
    }
}

template <class Vis>
__global__ __launch_bounds__(256) void k_walk_base(DevGraph g, Vis vis, EdgeTask* tasks, u64* ntasks, int64_t task_cap) {
    GC c = grid_coords();
    int64_t groups = ((int64_t)g.n + MF_WAVE - 1) / MF_WAVE;
    for (int64_t gi = c.wave; gi < groups; gi += c.nwaves) {
        int32_t v0 = (int32_t)(gi * MF_WAVE);
        int32_t v1 = v0 + MF_WAVE < g.n ? v0 + MF_WAVE : g.n;
        int64_t pb = g.nindex[v0], pe = g.nindex[v1];
        int64_t first_end = pe;
        if (pe - pb > MF_SPLIT) {
            first_end = pb + MF_SPLIT;
            int64_t extra = (pe - first_end + MF_SPLIT - 1) / MF_SPLIT;
            u64 t0 = 0;
            if (c.lane == 0) t0 = atomicAdd(ntasks, (u64)extra);
            t0 = __shfl((long long)t0, 0, MF_WAVE);
            for (int64_t k = c.lane; k < extra; k += MF_WAVE) {
                if ((int64_t)(t0 + k) < task_cap) {
                    tasks[t0 + k].v0 = v0;
                    tasks[t0 + k].p0 = first_end + k * MF_SPLIT;
                }
            }
        }
        walk_range(g, v0, v1, pb, first_end, vis);
    }
}

template <class Vis>
__global__ __launch_bounds__(256) void k_walk_tasks(DevGraph g, Vis vis, const EdgeTask* tasks, const u64* ntasks,
                                                    int64_t task_cap) {
    GC c = grid_coords();
    int64_t nt = (int64_t)*ntasks;
    if (nt > task_cap) nt = task_cap;
    for (int64_t t = c.wave; t < nt; t += c.nwaves) {
        int32_t v0 = tasks[t].v0;
        int32_t v1 = v0 + MF_WAVE < g.n ? v0 + MF_WAVE : g.n;
        int64_t pe = g.nindex[v1];
        int64_t p0 = tasks[t].p0;
        int64_t p1 = p0 + MF_SPLIT < pe ? p0 + MF_SPLIT : pe;
        walk_range(g, v0, v1, p0, p1, vis);
    }
}

template <class Vis>
__global__ __launch_bounds__(256) void k_walk_runs(DevGraph g, Vis vis) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < g.n; x += (int64_t)gridDim.x * blockDim.x) {
        int32_t len = g.ov_len[x];
        if (!len) continue;
        const int32_t* r = g.pool + g.ov_start[x];
        for (int32_t i = 0; i < len; i++) {
            int32_t y = r[i];
            if (y >= 0 && y != (int32_t)x) vis.visit((int32_t)x, y);
        }
    }
}

template <class Vis>
static void walk_all(GraphState& S, const Vis& vis) {
    const int B = S.blocks;
    k_fill<u64><<<1, 64, 0, S.stream>>>(S.d_walk_ntasks, 0ull, 1);
    k_walk_base<Vis><<<B, 256, 0, S.stream>>>(S.g, vis, (EdgeTask*)S.d_walk_tasks, S.d_walk_ntasks, S.walk_task_cap);
    k_walk_tasks<Vis><<<B, 256, 0, S.stream>>>(S.g, vis, (const EdgeTask*)S.d_walk_tasks, S.d_walk_ntasks, S.walk_task_cap);
    k_walk_runs<Vis><<<B, 256, 0, S.stream>>>(S.g, vis);
}

// ---------------------------------------------------------------- MIS
__global__ void k_finish_counts(uint32_t* mc, const uint32_t* cnt, int32_t n) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        mc[x] = (mc[x] & MF_MISBIT) | cnt[x];
}

hipError_t mf_set_mis(GraphState& S, const uint8_t* d_mis) {
    const int B = S.blocks;
    k_set_mis<<<B, 256, 0, S.stream>>>(S.g.mc, d_mis, S.g.n);
    k_fill<uint32_t><<<B, 256, 0, S.stream>>>(S.d_cnt, 0u, S.g.n);
    PassCountMis vis{S.g.mc, S.d_cnt, nullptr};
    walk_all(S, vis);
    k_finish_counts<<<B, 256, 0, S.stream>>>(S.g.mc, S.d_cnt, S.g.n);
    return hipGetLastError();
}

hipError_t mf_get_mis(GraphState& S, uint8_t* d_mis) {
    k_get_mis<<<S.blocks, 256, 0, S.stream>>>(S.g.mc, d_mis, S.g.n);
    return hipGetLastError();
}

hipError_t mf_flags(GraphState& S, const int32_t* d_u, const int32_t* d_v, int64_t count, uint8_t* d_ins) {
    k_flags<<<S.blocks, 256, 0, S.stream>>>(S.g, d_u, d_v, count, d_ins);
    return hipGetLastError();
}

// ---------------------------------------------------------------- check
__global__ void k_check_finish(const uint32_t* mc, const uint32_t* cnt, int32_t n, u64* acc) {
    u64 unc = 0, bad = 0;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = mc[x];
        uint32_t k = cnt[x];
        if (!(m & MF_MISBIT) && k == 0) unc++;
        if ((m & MF_CMASK) != k) bad++;
    }
    if (unc) atomicAdd(&acc[1], unc);
    if (bad) atomicAdd(&acc[2], bad);
}

// acc[3] = 1 when the edge walk's split tasks outgrew their list (some base
// entries were then not visited and the other three results are invalid).
__global__ void k_walk_overflow(const u64* ntasks, int64_t cap, u64* acc) {
    if (threadIdx.x == 0 && (int64_t)*ntasks > cap) acc[3] = 1;
}

hipError_t mf_check(GraphState& S, u64* d_acc) {
    const int B = S.blocks;
    k_fill<uint32_t><<<B, 256, 0, S.stream>>>(S.d_cnt, 0u, S.g.n);
    k_fill<u64><<<1, 64, 0, S.stream>>>(d_acc, 0ull, 4);
    PassCountMis vis{S.g.mc, S.d_cnt, d_acc};
    walk_all(S, vis);
    k_check_finish<<<B, 256, 0, S.stream>>>(S.g.mc, S.d_cnt, S.g.n, d_acc);
    k_walk_overflow<<<1, 64, 0, S.stream>>>(S.d_walk_ntasks, S.walk_task_cap, d_acc);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scan
// Exclusive prefix sum of int64 values, three passes (block sums, scan of
// the block sums in one block, local rescan plus offset).
#define SCAN_T 256
#define SCAN_PER 8
#define SCAN_TILE (SCAN_T * SCAN_PER)

__device__ int64_t block_excl_scan(int64_t v, int64_t* sh, int64_t& total) {
    int t = threadIdx.x;
    int lane = t & 63, w = t >> 6;
    int64_t incl = wave_incl_scan(v);
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int i = 0; i < SCAN_T / 64; i++) {
            int64_t s = sh[i];
            sh[i] = run;
            run += s;
        }
        sh[SCAN_T / 64] = run;
    }
    __syncthreads();
    int64_t r = sh[w] + incl - v;
    total = sh[SCAN_T / 64];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int64_t* in, int64_t count, int64_t* part) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int64_t s = 0;
    for (int i = 0; i < SCAN_PER; i++)
        if (base + i < count) s += in[base + i];
    int64_t tot;
    block_excl_scan(s, sh, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_parts(int64_t* part, int64_t nparts) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nparts; b0 += SCAN_T) {
        int64_t i = b0 + threadIdx.x;
        int64_t v = i < nparts ? part[i] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan(v, sh, tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_final(const int64_t* in, int64_t count, const int64_t* part,
                                                       int64_t* out, int64_t* total_out) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int64_t vals[SCAN_PER];
    int64_t s = 0;
    for (int i = 0; i < SCAN_PER; i++) {
        vals[i] = base + i < count ? in[base + i] : 0;
        s += vals[i];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(s, sh, tot) + part[blockIdx.x];
    for (int i = 0; i < SCAN_PER; i++) {
        if (base + i < count) out[base + i] = ex;
        ex += vals[i];
    }
    if (base < count && base + SCAN_PER >= count && total_out) *total_out = ex;
}

// out[i] = sum in[0..i); out[count] = total.  part needs ceil(count/2048).
void mf_scan_excl(const int64_t* in, int64_t count, int64_t* out, int64_t* part, hipStream_t st) {
    int64_t nb = (count + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 0) {
        k_fill<int64_t><<<1, 64, 0, st>>>(out, 0, 1);
        return;
    }
    k_scan_reduce<<<(unsigned)nb, SCAN_T, 0, st>>>(in, count, part);
    k_scan_parts<<<1, SCAN_T, 0, st>>>(part, nb);
    k_scan_final<<<(unsigned)nb, SCAN_T, 0, st>>>(in, count, part, out, out + count);
}

// ---------------------------------------------------------------- compaction plan
// The overlay's per-row counts, deleted base entries (ov_tomb) and live
// overlay entries (ov_alive), scanned together in one three-pass scan of
// three columns: doff and toff, the entries of each kind in the rows before,
// the new row starts nidx = nindex - doff + toff (a new row is its old row
// minus its deletions plus its tail), each with its total at [n], and the
// rows with a tail listed in row order, each at the count of such rows
// before it (no atomics).  part: three columns of block sums.
#define MF_PLAN_COLS 3
__global__ __launch_bounds__(SCAN_T) void k_plan_reduce(DevGraph g, int64_t* part) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int64_t sa = 0, sb = 0, sc = 0;
    for (int i = 0; i < SCAN_PER; i++)
        if (base + i < g.n) {
            const int32_t b = g.ov_alive[base + i];
            sa += g.ov_tomb[base + i];
            sb += b;
            sc += b > 0 ? 1 : 0;
        }
    int64_t ta, tb, tc;
    block_excl_scan(sa, sh, ta);
    block_excl_scan(sb, sh, tb);
    block_excl_scan(sc, sh, tc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ta;
        part[gridDim.x + blockIdx.x] = tb;
        part[2 * (int64_t)gridDim.x + blockIdx.x] = tc;
    }
}

// the columns of block sums, one workgroup each
__global__ __launch_bounds__(SCAN_T) void k_plan_parts(int64_t* part, int64_t nb) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    int64_t* col = part + (int64_t)blockIdx.x * nb;
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? col[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan(v, sh, tot);
        if (i < nb) col[i] = carry + ex;
        carry += tot;
    }
}

// A thread scans SCAN_PER consecutive rows; the offsets are read and the
// three output columns written through LDS in the other order (row
// base + i * SCAN_T + thread), so every load and store of a wave covers one
// contiguous span instead of 64 scattered lines.
__global__ __launch_bounds__(SCAN_T) void k_plan_final(DevGraph g, const int64_t* part, int64_t* doff, int64_t* toff,
                                                       int64_t* nidx, int32_t* trows, u64* ntrows) {
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    __shared__ int64_t tr[SCAN_TILE];
    const int64_t n = g.n;
    const int t = threadIdx.x;
    const int64_t blk = (int64_t)blockIdx.x * SCAN_TILE;
    const int64_t base = blk + (int64_t)t * SCAN_PER;
    for (int i = 0; i < SCAN_PER; i++) {
        const int64_t x = blk + i * SCAN_T + t;
        tr[i * SCAN_T + t] = x < n ? g.nindex[x] : 0;
    }
    int32_t a[SCAN_PER], b[SCAN_PER];
    int64_t sa = 0, sb = 0, sc = 0;
    for (int i = 0; i < SCAN_PER; i++) {
        a[i] = base + i < n ? g.ov_tomb[base + i] : 0;
        b[i] = base + i < n ? g.ov_alive[base + i] : 0;
        sa += a[i];
        sb += b[i];
        sc += b[i] > 0 ? 1 : 0;
    }
    int64_t ta, tb, tc;
    int64_t ea = block_excl_scan(sa, sh, ta) + part[blockIdx.x];  // its barriers also order tr
    int64_t eb = block_excl_scan(sb, sh, tb) + part[gridDim.x + blockIdx.x];
    int64_t ec = block_excl_scan(sc, sh, tc) + part[2 * (int64_t)gridDim.x + blockIdx.x];
    int64_t vd[SCAN_PER], vt[SCAN_PER], vn[SCAN_PER];
    for (int i = 0; i < SCAN_PER; i++) {
        vd[i] = ea;
        vt[i] = eb;
        vn[i] = tr[t * SCAN_PER + i] - ea + eb;
        if (base + i < n && b[i] > 0) trows[ec++] = (int32_t)(base + i);
        ea += a[i];
        eb += b[i];
    }
    if (base < n && base + SCAN_PER >= n) {
        doff[n] = ea;
        toff[n] = eb;
        nidx[n] = g.nindex[n] - ea + eb;
        *ntrows = (u64)ec;
    }
    int64_t* const out[3] = {doff, toff, nidx};
    for (int k = 0; k < 3; k++) {
        __syncthreads();  // tr is free
        for (int i = 0; i < SCAN_PER; i++) tr[t * SCAN_PER + i] = k == 0 ? vd[i] : k == 1 ? vt[i] : vn[i];
        __syncthreads();
        for (int i = 0; i < SCAN_PER; i++) {
            const int64_t x = blk + i * SCAN_T + t;
            if (x < n) out[k][x] = tr[i * SCAN_T + t];
        }
    }
}

hipError_t mf_compact_plan_counts(GraphState& S, int64_t* nidx, int64_t* doff, int64_t* toff, int32_t* trows,
                                  u64* ntrows, int64_t* part) {
    const int64_t nb = (S.g.n + SCAN_TILE - 1) / SCAN_TILE;
    k_plan_reduce<<<(unsigned)nb, SCAN_T, 0, S.stream>>>(S.g, part);
    k_plan_parts<<<MF_PLAN_COLS, SCAN_T, 0, S.stream>>>(part, nb);
    k_plan_final<<<(unsigned)nb, SCAN_T, 0, S.stream>>>(S.g, part, doff, toff, nidx, trows, ntrows);
    return hipGetLastError();
}

// ---------------------------------------------------------------- compaction
// New row lengths: old length - deleted base entries + live overlay entries.
__global__ __launch_bounds__(256) void k_newdeg(DevGraph g, int64_t* deg) {
    for (int64_t x = (int64_t)blockIdx.x * MF_BLOCK + threadIdx.x; x < g.n; x += (int64_t)gridDim.x * MF_BLOCK)
        deg[x] = (g.nindex[x + 1] - g.nindex[x]) - g.ov_tomb[x] + g.ov_alive[x];
}

__device__ __forceinline__ u64 below_mask(int64_t p) { return (1ull << (p & 63)) - 1ull; }

// A deletion-mark word read as three 32-bit words: the mask and the low half
// of pre.  Differences of pre inside one row are below 2^31, so they come out
// exact in 32-bit arithmetic.
struct TombWord3 {
    u64 mask;
    uint32_t pre;
};
__device__ __forceinline__ TombWord3 load_tw3(const TombWord* t) {
    const uint3 w = *reinterpret_cast<const uint3*>(t);
    return TombWord3{(u64)w.x | ((u64)w.y << 32), w.z};
}
// deletions of the row (whose first deletion-mark count is d0) before entry p, from the word of p
__device__ __forceinline__ int32_t tomb_before(const TombWord3& t, int64_t p, int64_t d0) {
    return (int32_t)(t.pre + (uint32_t)__popcll(t.mask & ((1ull << (p & 63)) - 1ull)) - (uint32_t)d0);
}

// deleted entries of the array before entry p
__device__ __forceinline__ int64_t tomb_rank(const TombWord* tw, int64_t p) {
    const TombWord t = tw[p >> 6];
    return (int64_t)t.pre + __popcll(t.mask & below_mask(p));
}

// Deletion marks of the old rows in position order (a deleted base edge is
// stored as ~v): one wave per counting chunk, a ballot per word, `pre`
// counted from the chunk's start (k_words_fix adds the chunks before), and
// the chunk's count.  Words are loaded 16 at a time so each wave keeps
// 16 loads in flight.
#define MF_TW_UNROLL 16
__global__ __launch_bounds__(256) void k_tomb_words(const int32_t* nlist, int64_t m, TombWord* tw, int64_t nchunks,
                                                    int64_t* ccnt) {
    GC c = grid_coords();
    const int64_t nw = mf_tomb_words(m);
    if (m == 0) {  // one word, no entries
        if (c.tid == 0) {
            tw[0] = TombWord{0ull, 0ull};
            ccnt[0] = 0;
        }
        return;
    }
    for (int64_t t = c.wave; t < nchunks; t += c.nwaves) {
        const int64_t w0 = t * (MF_TCHUNK / 64);
        const int64_t w1 = w0 + MF_TCHUNK / 64 < nw ? w0 + MF_TCHUNK / 64 : nw;
        u64 run = 0;
        for (int64_t w = w0; w < w1; w += MF_TW_UNROLL) {
            // every load in bounds (clamped) and unconditional, so all of
            // them are in flight before the first ballot
            int32_t y[MF_TW_UNROLL];
#pragma unroll
            for (int j = 0; j < MF_TW_UNROLL; j++) {
                const int64_t p = (w + j) * 64 + c.lane;
                y[j] = nlist[p < m ? p : m - 1];
            }
#pragma unroll
            for (int j = 0; j < MF_TW_UNROLL; j++) {
                const int64_t p = (w + j) * 64 + c.lane;
                const u64 mask = __ballot(y[j] < 0 && p < m && w + j < w1);
                if (c.lane == 0 && w + j < w1) tw[w + j] = TombWord{mask, run};
                run += __popcll(mask);
            }
        }
        if (c.lane == 0) ccnt[t] = (int64_t)run;
    }
}

__global__ void k_words_fix(TombWord* tw, int64_t nw, const int64_t* cpre) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
        tw[i].pre += (u64)cpre[i / (MF_TCHUNK / 64)];
}

// As k_words_fix for marks counted per merge-path chunk: a wave per chunk
// adds the deletions of the chunks before it to its words.
__global__ __launch_bounds__(256) void k_words_fix_path(TombWord* tw, int64_t nw, const int64_t* cst, int64_t nch,
                                                        const int64_t* cpre) {
    GC c = grid_coords();
    for (int64_t t = c.wave; t < nch; t += c.nwaves) {
        const int64_t w0 = cst[t] >> 6, w1 = t + 1 < nch ? cst[t + 1] >> 6 : nw;
        const u64 add = (u64)cpre[t];
        for (int64_t w = w0 + c.lane; w < w1; w += MF_WAVE) tw[w].pre += add;
    }
}

// Deletion marks of the old rows in position order, counted per MF_TCHUNK entries.
static void tomb_marks(GraphState& S, TombWord* out, const MergeLists& L) {
    const int64_t m = S.g.m;
    const int64_t nch = mf_tomb_chunks(m);
    const int B = (int)std::min<int64_t>(S.blocks, (nch + 3) / 4);
    k_tomb_words<<<B, 256, 0, S.stream>>>(S.g.nlist, m, out, nch, L.ccnt);
    mf_scan_excl(L.ccnt, nch, L.cpre, L.cpart, S.stream);
    k_words_fix<<<S.blocks, 256, 0, S.stream>>>(out, mf_tomb_words(m), L.cpre);
}

// Every kept base entry to its compacted position: row x starts at nidx[x],
// entry p keeps its place minus the deletions before it in the row, i.e.
// dst = nidx[x] + doff[x] + (p - start) - (deletions before p in the array).
// Balanced by entries over all rows, hub rows spread over many waves.
//
// Sorted (the old rows have a sorted copy): the same walk also writes the
// deletion marks of the old sorted copy, since entry p of the sorted copy
// lies in the same row as entry p of the rows: it is deleted when the entry
// at its row position spos[p] is (bit of the position marks tw).  Each
// group's ballot goes into the aligned word it covers (64 contiguous entries
// straddle at most one word boundary), every word of the chunk is written
// with `pre` counted from the chunk's start, and the chunk's count goes to
// ccnt (zeroed before: chunks of empty rows are not visited); merge-path
// chunk starts are multiples of 64 entries, so every word belongs to one
// chunk, and the last chunk also owns the word past the end.
template <bool Sorted>
__global__ __launch_bounds__(256) void k_copy_base(DevGraph g, const int64_t* nidx, const int64_t* doff,
                                                   const TombWord* tw, const int64_t* cst, int64_t nch,
                                                   int32_t* out, TombWord* tws, int64_t* ccnt) {
    // Entries come in order within a chunk, so the deletions before an
    // entry are counted as the walk goes (a ballot per group), from the
    // chunk's start (a multiple of 64 entries: its word's `pre`).
    const int lane = lane_id();
    const int64_t m = g.m;
    const int64_t nw = mf_tomb_words(m);
    int64_t chunk = -1;
    u64 grun = 0;
    // sorted-order marks of the wave's current chunk
    int64_t wcur = 0, wend = 0;
    u64 mcur = 0, run = 0;
    auto close_chunk = [&]() {
        if (chunk < 0) return;
        for (; wcur < wend; wcur++) {  // the last, partial word and any word past the end
            if (lane == 0) tws[wcur] = TombWord{mcur, run};
            run += __popcll(mcur);
            mcur = 0;
        }
        if (lane == 0) ccnt[chunk] = (int64_t)run;
    };
    struct D {
        int32_t y[MF_EW_GROUPS];
        int64_t a[MF_EW_GROUPS];   // Sorted: the sorted entry's row position
        u64 word[MF_EW_GROUPS];    // ... and its position-marks word
    };
    entry_walk4<PayI64>(
        g.nindex, 0, g.n, cst, nch, [&](int64_t x) { return PayI64{nidx[x] + doff[x]}; },
        [&](const WalkStep<PayI64>& st) {
            D d;
#pragma unroll
            for (int j = 0; j < MF_EW_GROUPS; j++) {
                d.y[j] = g.nlist[st.q(j)];
                if (Sorted) d.a[j] = st.start(j) + (int64_t)g.spos[st.q(j)];
            }
            return d;
        },
        [&](const WalkStep<PayI64>&, D& d) {
            if (Sorted) {
                // clamped: a lane past the step's end may pair its row with another row's entry
#pragma unroll
                for (int j = 0; j < MF_EW_GROUPS; j++) d.word[j] = tw[d.a[j] >> 6 < nw ? d.a[j] >> 6 : nw - 1].mask;
            }
        },
        [&](const WalkStep<PayI64>& st, const D& d) {
            if (st.chunk != chunk) {  // wave-uniform
                if (Sorted) close_chunk();
                chunk = st.chunk;
                grun = tw[cst[chunk] >> 6].pre;
                if (Sorted) {
                    wcur = cst[chunk] >> 6;
                    wend = chunk + 1 < nch ? cst[chunk + 1] >> 6 : nw;
                    mcur = 0;
                    run = 0;
                }
            }
#pragma unroll
            for (int j = 0; j < MF_EW_GROUPS; j++) {
                const bool valid = st.valid(j);
                const u64 bal = __ballot(valid && d.y[j] < 0);
                if (valid && d.y[j] >= 0) {
                    const RowRec<PayI64>& r = st.tab[st.rl[j]];
                    out[r.pay.v + (st.pp(j) - r.start) - (int64_t)(grun + __popcll(bal & below_mask(lane)))] = d.y[j];
                }
                grun += __popcll(bal);
                if (Sorted) {
                    const int cnt = __popcll(__ballot(valid));
                    if (cnt) {  // groups past the window's end hold no entry (wave-uniform)
                        const u64 sb = __ballot(valid && ((d.word[j] >> (d.a[j] & 63)) & 1ull));
                        const int sh = (int)((st.p0 + MF_WAVE * j) & 63);  // the group's first entry's bit
                        mcur |= sb << sh;
                        if (sh + cnt >= 64) {  // the word is complete
                            if (lane == 0) tws[wcur] = TombWord{mcur, run};
                            run += __popcll(mcur);
                            wcur++;
                            mcur = sh ? (sb >> (64 - sh)) : 0ull;
                        }
                    }
                }
            }
        });
    if (Sorted) {
        close_chunk();
        // the last chunk holds no entry (m a multiple of 64, its only word is
        // the one past the end): written here
        if (blockIdx.x == 0 && threadIdx.x == 0 && cst[nch - 1] == cst[nch])
            for (int64_t w = cst[nch - 1] >> 6; w < nw; w++) tws[w] = TombWord{0ull, 0ull};
    }
}

// live overlay entries follow the base part, in insertion order; with
// L.tval, they also go to the row's slice of the tail lists with their
// insertion index.  With check_order, a row whose entries are not strictly
// ascending after its tail was appended sets flags |= 1 (its kept part is
// ascending when the old rows were, so only the tail and its first entry's
// step need testing).
#define MF_CR_ROWS 4  // rows per thread and step: each level of loads is issued for all of them together
__global__ __launch_bounds__(256) void k_copy_runs(DevGraph g, const int64_t* nidx, int32_t* out, MergeLists L,
                                                   int check_order, int* flags) {
    bool unsorted = false;
    const int64_t stride = (int64_t)gridDim.x * MF_BLOCK;
    for (int64_t x0 = (int64_t)blockIdx.x * MF_BLOCK + threadIdx.x; x0 < g.n; x0 += MF_CR_ROWS * stride) {
        int64_t x[MF_CR_ROWS];
        int32_t len[MF_CR_ROWS];
#pragma unroll
        for (int k = 0; k < MF_CR_ROWS; k++) {
            x[k] = x0 + k * stride;
            len[k] = g.ov_len[x[k] < g.n ? x[k] : g.n - 1];  // clamped, then dropped below
            if (x[k] >= g.n) len[k] = 0;
        }
        int64_t row0[MF_CR_ROWS], dst[MF_CR_ROWS], t0[MF_CR_ROWS], st[MF_CR_ROWS];
#pragma unroll
        for (int k = 0; k < MF_CR_ROWS; k++) {
            if (!len[k]) continue;
            row0[k] = nidx[x[k]];
            dst[k] = row0[k] + (g.nindex[x[k] + 1] - g.nindex[x[k]]) - g.ov_tomb[x[k]];
            t0[k] = L.tval ? L.toff[x[k]] : 0;
            st[k] = g.ov_start[x[k]];
        }
#pragma unroll
        for (int k = 0; k < MF_CR_ROWS; k++) {
            if (!len[k]) continue;
            int64_t d = dst[k];
            int32_t prev = (check_order && d > row0[k]) ? out[d - 1] : -1;
            int32_t j = 0;
            const int32_t* r = g.pool + st[k];
            for (int32_t i = 0; i < len[k]; i++) {
                const int32_t y = r[i];
                if (y < 0) continue;
                out[d++] = y;
                if (y <= prev) unsorted = true;
                prev = y;
                if (L.tval) {
                    L.tval[t0[k] + j] = y;
                    L.tidx[t0[k] + j] = j;
                }
                j++;
            }
        }
    }
    if (check_order && unsorted) atomicOr(flags, 1);
}

hipError_t mf_compact_plan(GraphState& S, int64_t* d_newidx) {
    k_newdeg<<<S.blocks, 256, 0, S.stream>>>(S.g, S.d_deg);
    mf_scan_excl(S.d_deg, S.g.n, d_newidx, S.d_scan_part, S.stream);
    return hipGetLastError();
}

hipError_t mf_compact_copy(GraphState& S, const int64_t* d_newidx, int32_t* d_out, const MergeLists& L, int* d_flags) {
    merge_path(S.g.nindex, 0, S.g.n, L.ncst, L.cst, S.stream);
    tomb_marks(S, L.tw, L);
    if (S.g.sval) {  // the sorted copy's marks in the same walk, counted per merge-path chunk
        k_fill<int64_t><<<1024, 256, 0, S.stream>>>(L.ccnt, 0, L.ncst);
        k_copy_base<true><<<S.blocks, 256, 0, S.stream>>>(S.g, d_newidx, L.doff, L.tw, L.cst, L.ncst, d_out, L.tws,
                                                          L.ccnt);
    } else {
        k_copy_base<false><<<S.blocks, 256, 0, S.stream>>>(S.g, d_newidx, L.doff, L.tw, L.cst, L.ncst, d_out, nullptr,
                                                           nullptr);
    }
    k_fill<int><<<1, 64, 0, S.stream>>>(d_flags, S.g.sorted ? 0 : 1, 1);
    k_copy_runs<<<S.blocks, 256, 0, S.stream>>>(S.g, d_newidx, d_out, L, S.g.sorted ? 1 : 0, d_flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------- bytes
// V1 = S ∪ N(S) (S = endpoints of batch b), V2 = V1 ∪ N(V1).
__global__ __launch_bounds__(256) void k_mark_v1(DevGraph g, const int32_t* seeds, const u64* nseeds, uint32_t* mark,
                                                 uint32_t v1tag) {
    GC c = grid_coords();
    int64_t ns = (int64_t)*nseeds;
    for (int64_t i = c.wave; i < ns; i += c.nwaves) {
        int32_t s = seeds[i];
        if (c.lane == 0) mark[s] = v1tag;
        Row r = get_row(g, s);
        int64_t L = r.len();
        for (int64_t j = c.lane; j < L; j += MF_WAVE) {
            int32_t y = r.at(j);
            if (y >= 0) mark[y] = v1tag;
        }
    }
}

__global__ void k_count_v(DevGraph g, const uint32_t* mark, uint32_t v1tag, uint32_t v2tag, u64* acc) {
    u64 n1 = 0, n2 = 0, deg = 0;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < g.n; x += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = mark[x];
        if (m == v1tag) {
            n1++;
            n2++;
            deg += (u64)((g.nindex[x + 1] - g.nindex[x]) - g.ov_tomb[x] + g.ov_alive[x]);
        } else if (m == v2tag) {
            n2++;
        }
    }
    if (n1) atomicAdd(&acc[0], n1);
    if (n2) atomicAdd(&acc[1], n2);
    if (deg) atomicAdd(&acc[2], deg);
}

hipError_t mf_batch_bytes(GraphState& S, const int32_t* d_seeds, const u64* d_nseeds, uint32_t t, u64* d_acc) {
    const int B = S.blocks;
    uint32_t v1tag = 2 * t + 1, v2tag = 2 * t;
    k_fill<u64><<<1, 64, 0, S.stream>>>(d_acc, 0ull, 4);
    k_mark_v1<<<B, 256, 0, S.stream>>>(S.g, d_seeds, d_nseeds, S.d_bmark, v1tag);
    PassMarkV2 vis{S.d_bmark, v1tag, v2tag};
    walk_all(S, vis);
    k_count_v<<<B, 256, 0, S.stream>>>(S.g, S.d_bmark, v1tag, v2tag, d_acc);
    return hipGetLastError();
}

// ---------------------------------------------------------------- misc
hipError_t mf_widen(const int32_t* d_in, int64_t* d_out, int64_t count, hipStream_t st) {
    k_widen32<<<1024, 256, 0, st>>>(d_in, d_out, count);
    return hipGetLastError();
}

hipError_t mf_validate_csr(GraphState& S, int* d_flags) {
    k_fill<int><<<1, 64, 0, S.stream>>>(d_flags, 0, 1);
    k_validate_offsets<<<S.blocks, 256, 0, S.stream>>>(S.g.nindex, S.g.n, S.g.m, d_flags);
    // entries are only inspected when the offsets are sane (host checks flag 8 first)
    return hipGetLastError();
}

hipError_t mf_validate_entries(GraphState& S, int* d_flags) {
    k_validate_entries<<<S.blocks, 256, 0, S.stream>>>(S.g.nindex, S.g.nlist, S.g.n, S.g.m, d_flags);
    return hipGetLastError();
}

void mf_fill_u32(uint32_t* p, uint32_t v, int64_t count, hipStream_t st) {
    k_fill<uint32_t><<<1024, 256, 0, st>>>(p, v, count);
}
void mf_fill_i32(int32_t* p, int32_t v, int64_t count, hipStream_t st) {
    k_fill<int32_t><<<1024, 256, 0, st>>>(p, v, count);
}
void mf_fill_u64(u64* p, u64 v, int64_t count, hipStream_t st) { k_fill<u64><<<1024, 256, 0, st>>>(p, v, count); }
void mf_fill_i64(int64_t* p, int64_t v, int64_t count, hipStream_t st) {
    k_fill<int64_t><<<1024, 256, 0, st>>>(p, v, count);
}

// ---------------------------------------------------------------- row index
// Rows that are not ascending (an input in file order, or base rows after a
// compaction: base order, then insertions in insertion order) keep their
// order, which the rule observes; a sorted copy of every row, with each
// entry's position in its row, lets adjacency tests binary-search anyway.
// Built with rocPRIM's segmented radix sort over row ranges of at most 2^30
// entries; checked for repeated neighbours.
#include <hipcub/device/device_segmented_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace {
struct RebaseOffset {
    int64_t base;
    __host__ __device__ int operator()(const int64_t& x) const { return (int)(x - base); }
};
}  // namespace

// Position of every entry of rows [r0, r1) inside its row, at pos[p - first entry].
__global__ __launch_bounds__(256) void k_row_iota(const int64_t* nindex, int32_t r0, int32_t r1, int32_t* pos) {
    const int64_t base = nindex[r0];
    entry_walk(
        nindex, r0, r1, [](int64_t) { return (int64_t)0; },
        [&](int64_t p, int64_t, int64_t start, int64_t) { pos[p - base] = (int32_t)(p - start); });
}

// Checks a sorted row copy against its rows, entry by entry: every row of
// sval strictly ascending (flags |= 16 on a value repeated, i.e. a row that
// lists a neighbour twice) and every sval[p] found at its spos[p] in the row
// (flags |= 32 otherwise: the copy is wrong).  Strictly ascending values that
// all occur in the row, as many as the row has, make spos a permutation.
__global__ __launch_bounds__(256) void k_check_row_index(const int64_t* nindex, const int32_t* nlist,
                                                         const int32_t* sval, const int32_t* spos, int32_t n,
                                                         int* flags) {
    int f = 0;
    entry_walk(
        nindex, 0, n, [&](int64_t x) { return nindex[x + 1]; },
        [&](int64_t p, int64_t, int64_t start, int64_t end) {
            int32_t v = sval[p];
            int32_t q = spos[p];
            if (q < 0 || q >= end - start || nlist[start + q] != v) f |= 32;
            if (p > start) {
                int32_t pv = sval[p - 1];
                if (pv == v) f |= 16;
                else if (pv > v) f |= 32;
            }
        });
    if (f) atomicOr(flags, f);
}

// ---------------------------------------------------------------- merged row copy
// After a compaction row x is K ++ T: the old row's kept entries in old
// order, then its live overlay entries in insertion order.  Its sorted copy
// merges the old row's sorted order (the row itself when rows were
// ascending, else the old sorted copy), minus the deleted entries, with the
// sorted tail:
//   a kept entry of old sorted rank k, value v, old position q lands at
//     k - #{deleted values < v} + #{tail values < v}, new position
//     q - #{deleted positions < q};
//   the tail entry of sorted rank j, value w, insertion index i lands at
//     #{old values < w} - #{deleted values < w} + j, new position |K| + i.
// Tail values never occur in the old row (re-inserting a deleted base edge
// restores it in place), so all values of the new row are distinct ranks.
// The deletion counts come from the deletion marks: in position order for
// #{deleted positions < q}, in sorted order for #{deleted values < v} (the
// same marks when rows were ascending), each one load and a popcount.
struct OldRows {
    const int64_t* nindex;
    const int32_t* nlist;  // with deletions as ~v
    const int32_t* sval;   // sorted copy, or null when rows are ascending
    const int32_t* spos;
};

__device__ __forceinline__ int32_t old_sorted_val(const OldRows& o, int64_t p) {
    return o.sval ? o.sval[p] : decode(o.nlist[p]);
}

// first index in [lo, hi) with a[i] >= v
__device__ __forceinline__ int64_t lower32(const int32_t* a, int64_t lo, int64_t hi, int32_t v) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Wave shift by one lane: lane l gets v of lane l - 1, lane 0 gets 0 (DPP).
__device__ __forceinline__ int wave_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }

// The merge, one walk of the old sorted order (entry_walk4).  Each old entry
// k of a row, value v, kept or deleted, with t(k) = #{tail values < v}:
//   - a kept entry lands at (kept entries before k) + t(k);
//   - it places the tail entries of sorted rank j in [t(k-1), t(k)) (values
//     between the old values k-1 and k) at (kept entries before k) + j, and
//     the row's last old entry those of rank j >= t(k) at (kept entries of
//     the row) + j; rows with no old entry get their tails in
//     k_merge_tail_empty.
// t(k-1) comes from the lane before (a wave shift), from the group before,
// from the step before when it ended at k-1, or, at a chunk's start inside a
// row, from the old value k-1 itself.  A row's values come from the walk's
// window table, with its first two sorted tail values and their insertion
// indices; longer tails are searched.  HasSorted: the old rows had a sorted
// copy (else they were ascending and are their own sorted order).
struct PayMerge {
    int64_t d0, t0, dst;  // deletions before the row; its tail's offset; its new start
    int64_t end;          // the old row's end
    int32_t tn, kr;       // tail entries; kept old entries
    int32_t tv0, tv1;     // the first two sorted tail values (INT32_MAX past the tail)
    int32_t ti0, ti1;     // ... and their insertion indices
};

template <bool HasSorted>
__global__ __launch_bounds__(256) void k_merge_base(OldRows o, int32_t n, MergeLists L, const int32_t* tsv,
                                                    const int32_t* tsi, int64_t ntail, const int64_t* nidx,
                                                    int32_t* nsval, int32_t* nspos) {
    const TombWord* tw = L.tw;
    const TombWord* tws = HasSorted ? L.tws : L.tw;
    const int64_t nw = mf_tomb_words(o.nindex[n]);
    const int lane = lane_id();
    int64_t carry_p = -2;  // wave-uniform: the last entry of the step stored before, and its t
    int carry_t = 0;
    // per group: the entry's value and old position; the sorted-order marks'
    // word, folded into the row's deletions before the entry (ds) as soon as
    // load2 runs; the position-order marks' word of its old position
    struct D {
        int32_t v[MF_EW_GROUPS], q[MF_EW_GROUPS], ds[MF_EW_GROUPS];
        TombWord3 ws[MF_EW_GROUPS], wp[MF_EW_GROUPS];
    };
    // #{tail values < v} in a row
    auto tails_below = [&](const PayMerge& r, int32_t v) {
        if (r.tn <= 2) return (r.tv0 < v ? 1 : 0) + (r.tv1 < v ? 1 : 0);
        return (int)(lower32(tsv, r.t0, r.t0 + r.tn, v) - r.t0);
    };
    entry_walk4<PayMerge>(
        o.nindex, 0, n, L.cst, L.ncst,
        [&](int64_t x) {
            const int64_t t0 = L.toff[x];
            const int32_t tn = (int32_t)(L.toff[x + 1] - t0);
            // a row's first two tail entries ride along with the row (clamped
            // loads: the tail lists hold at least one entry even when no row has a tail)
            const int64_t i0 = t0 < ntail ? t0 : 0, i1 = t0 + 1 < ntail ? t0 + 1 : 0;
            const int32_t a = tsv[i0], b = tsv[i1], ia = tsi[i0], ib = tsi[i1];
            const int64_t dst = nidx[x];
            return PayMerge{L.doff[x], t0, dst, o.nindex[x + 1], tn, (int32_t)(nidx[x + 1] - dst - tn),
                            tn >= 1 ? a : INT32_MAX, tn >= 2 ? b : INT32_MAX, ia, ib};
        },
        [&](const WalkStep<PayMerge>& st) {
            D d;
#pragma unroll
            for (int j = 0; j < MF_EW_GROUPS; j++) {
                const int64_t q = st.q(j);
                if (HasSorted) {
                    d.q[j] = o.spos[q];
                    d.v[j] = o.sval[q];
                } else {
                    d.v[j] = o.nlist[q];
                    d.q[j] = (int32_t)(q - st.start(j));
                }
                d.ws[j] = load_tw3(&tws[q >> 6]);
            }
            return d;
        },
        [&](const WalkStep<PayMerge>& st, D& d) {
#pragma unroll
            for (int j = 0; j < MF_EW_GROUPS; j++) d.ds[j] = tomb_before(d.ws[j], st.pp(j), st.tab[st.rl[j]].pay.d0);
            if (HasSorted) {
                // clamped: a lane past the step's end may pair its row with another row's entry
#pragma unroll
                for (int j = 0; j < MF_EW_GROUPS; j++) {
                    const int64_t w = (st.start(j) + d.q[j]) >> 6;
                    d.wp[j] = load_tw3(&tw[w < nw ? w : nw - 1]);
                }
            }
        },
        [&](const WalkStep<PayMerge>& st, const D& d) {
            // t of the entry before the step's first one (wave-uniform, used by lane 0)
            int tprev0;
            if (carry_p == st.p0 - 1) {
                tprev0 = carry_t;
            } else {  // a chunk's first step: from the old value before it, when the row has one
                tprev0 = 0;
                const int64_t s0 = read_lane(st.start(0), 0);
                if (st.p0 > s0) {
                    const PayMerge r0 = st.tab[__builtin_amdgcn_readfirstlane(st.rl[0])].pay;
                    tprev0 = __builtin_amdgcn_readfirstlane(tails_below(r0, old_sorted_val(o, st.p0 - 1)));
                }
            }
            int tk[MF_EW_GROUPS];
#pragma unroll
            for (int j = 0; j < MF_EW_GROUPS; j++) {
                const int64_t pp = st.pp(j);
                const bool valid = pp < st.lim;
                const RowRec<PayMerge>& rr = st.tab[st.rl[j]];  // LDS
                const int64_t start = rr.start, d0 = rr.pay.d0, t0 = rr.pay.t0, dst = rr.pay.dst;
                const int32_t tn = rr.pay.tn;
                int32_t v = d.v[j];
                bool deleted;
                int64_t dpos = 0;
                const int64_t dsort = d.ds[j];
                if (HasSorted) {
                    const int64_t a = start + d.q[j];
                    deleted = (d.wp[j].mask >> (a & 63)) & 1ull;
                    dpos = tomb_before(d.wp[j], a, d0);
                } else {
                    deleted = v < 0;
                    if (deleted) v = ~v;
                    dpos = dsort;
                }
                const int32_t tv0 = rr.pay.tv0, tv1 = rr.pay.tv1;
                tk[j] = !valid ? 0
                        : tn <= 2 ? (tv0 < v ? 1 : 0) + (tv1 < v ? 1 : 0)
                                  : (int)(lower32(tsv, t0, t0 + tn, v) - t0);
                // t of the entry before this lane's
                const int tfirst = j == 0 ? tprev0 : __builtin_amdgcn_readlane(tk[j - 1], MF_WAVE - 1);
                int tp = wave_shr1(tk[j]);
                if (lane == 0) tp = tfirst;
                if (pp == start) tp = 0;
                if (!valid) continue;
                const int64_t kb = (pp - start) - dsort;  // kept entries before this one
                if (!deleted) {
                    const int64_t at = kb + tk[j];
                    nsval[dst + at] = v;
                    nspos[dst + at] = d.q[j] - (int32_t)dpos;
                }
                const int hi = pp + 1 == rr.pay.end ? tn : tk[j];
                if (tp < hi) {
                    const int32_t kr = rr.pay.kr, ti0 = rr.pay.ti0, ti1 = rr.pay.ti1;
                    for (int t = tp; t < hi; t++) {
                        const int64_t at = t < tk[j] ? kb + t : kb + (deleted ? 0 : 1) + t;
                        int32_t w = t == 0 ? tv0 : tv1, wi = t == 0 ? ti0 : ti1;  // values, not addresses
                        if (tn > 2) {
                            w = tsv[t0 + t];
                            wi = tsi[t0 + t];
                        }
                        nsval[dst + at] = w;
                        nspos[dst + at] = kr + wi;
                    }
                }
            }
            // the step's last entry, for the next step
            const int64_t last = (st.lim < st.p0 + MF_WAVE * MF_EW_GROUPS ? st.lim : st.p0 + MF_WAVE * MF_EW_GROUPS) - 1;
            const int gl = (int)((last - st.p0) >> 6), ll = (int)((last - st.p0) & 63);
            const int tl = gl == 0 ? tk[0] : gl == 1 ? tk[1] : gl == 2 ? tk[2] : tk[3];
            carry_t = __builtin_amdgcn_readlane(tl, ll);
            carry_p = last;
        });
}

// Tail entries of rows with no old entry: the whole new row, in sorted order
// (a coalesced pass over every row's offsets).
__global__ __launch_bounds__(256) void k_merge_tail_empty(const int64_t* nindex, int32_t n, MergeLists L,
                                                          const int32_t* tsv, const int32_t* tsi, const int64_t* nidx,
                                                          int32_t* nsval, int32_t* nspos) {
    GC c = grid_coords();
#pragma unroll 4
    for (int64_t x = c.tid; x < n; x += c.nthreads) {
        const int64_t t0 = L.toff[x], t1 = L.toff[x + 1], b = nindex[x], e = nindex[x + 1];
        if (t1 == t0 || e != b) continue;
        const int64_t dst = nidx[x];
        for (int64_t j = 0; j < t1 - t0; j++) {
            nsval[dst + j] = tsv[t0 + j];
            nspos[dst + j] = tsi[t0 + j];
        }
    }
}

// Sorts the per-row lists of the rows listed in rows[0, *nrows) by rank
// (values are distinct within a row: each entry goes to the number of
// smaller ones).  A wave takes 64 listed rows at a time: a list of at most
// MF_SORT_THREAD entries is ranked by its lane alone, one of at most
// MF_SORT_WAVE by the whole wave, longer ones are listed in big[] for the
// library's segmented sort.
#define MF_SORT_THREAD 8
#define MF_SORT_WAVE 256
__global__ __launch_bounds__(256) void k_sort_small(const int64_t* off, const int32_t* rows, const u64* nrows,
                                                    const int32_t* kin, const int32_t* vin, int32_t* kout,
                                                    int32_t* vout, int32_t* big, u64* nbig) {
    GC c = grid_coords();
    const int64_t nr = (int64_t)*nrows;
    for (int64_t base = c.wave * MF_WAVE; base < nr; base += c.nwaves * MF_WAVE) {
        const int64_t i = base + c.lane;
        int64_t b = 0, L = 0;
        int32_t x = 0;
        if (i < nr) {
            x = rows[i];
            b = off[x];
            L = off[x + 1] - b;
        }
        if (L > MF_SORT_WAVE) big[atomicAdd(nbig, 1ull)] = x;
        if (L <= MF_SORT_THREAD) {
            for (int64_t a = 0; a < L; a++) {
                const int32_t v = kin[b + a];
                int64_t r = 0;
                for (int64_t j = 0; j < L; j++) r += kin[b + j] < v ? 1 : 0;
                kout[b + r] = v;
                if (vout) vout[b + r] = vin[b + a];
            }
        }
        u64 med = __ballot(L > MF_SORT_THREAD && L <= MF_SORT_WAVE);
        while (med) {
            const int src = __builtin_ctzll(med);
            med &= med - 1;
            const int64_t rb = __shfl(b, src, MF_WAVE);
            const int64_t rl = __shfl(L, src, MF_WAVE);
            for (int64_t i0 = 0; i0 < rl; i0 += MF_WAVE) {
                const int64_t me = i0 + c.lane;
                const int32_t v = me < rl ? kin[rb + me] : 0;
                int64_t rank = 0;
                for (int64_t j0 = 0; j0 < rl; j0 += MF_WAVE) {
                    const int32_t u = j0 + c.lane < rl ? kin[rb + j0 + c.lane] : INT32_MAX;
                    const int64_t lim = rl - j0 < MF_WAVE ? rl - j0 : MF_WAVE;
                    for (int t = 0; t < (int)lim; t++) rank += __shfl(u, t, MF_WAVE) < v ? 1 : 0;
                }
                if (me < rl) {
                    kout[rb + rank] = v;
                    if (vout) vout[rb + rank] = vin[rb + me];
                }
            }
        }
    }
}

hipError_t mf_build_row_index(GraphState& S, int32_t* sval, int32_t* spos, int* d_flags) {
    const DevGraph& g = S.g;
    hipStream_t st = S.stream;
    if (g.m == 0) return hipSuccess;
    std::vector<int64_t> ni;
    const int64_t chunk = int64_t(1) << 30;
    std::vector<int32_t> cuts{0};
    if (g.m > chunk) {  // row ranges of at most 2^30 entries
        ni.resize((size_t)g.n + 1);
        hipError_t e = hipMemcpyAsync(ni.data(), g.nindex, ni.size() * 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        int64_t start = 0;
        for (int32_t r = 0; r < g.n; r++) {
            if (ni[r + 1] - start > chunk && ni[r] > start) {
                cuts.push_back(r);
                start = ni[r];
            }
        }
    }
    cuts.push_back(g.n);
    int64_t max_items = g.m;
    if (!ni.empty()) {
        max_items = 0;
        for (size_t k = 0; k + 1 < cuts.size(); k++) max_items = std::max(max_items, ni[cuts[k + 1]] - ni[cuts[k]]);
    }
    if (max_items > (int64_t)INT32_MAX) return hipErrorInvalidValue;  // one row of more than 2^31 entries
    int32_t* pos_in = nullptr;
    hipError_t e = hipMalloc((void**)&pos_in, (size_t)max_items * 4);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    for (size_t k = 0; k + 1 < cuts.size() && e == hipSuccess; k++) {
        int32_t r0 = cuts[k], r1 = cuts[k + 1];
        int64_t b, en;
        if (ni.empty()) {
            b = 0;
            en = g.m;
        } else {
            b = ni[r0];
            en = ni[r1];
        }
        int items = (int)(en - b);
        if (items == 0) continue;
        k_row_iota<<<S.blocks, 256, 0, st>>>(g.nindex, r0, r1, pos_in);
        rocprim::transform_iterator<const int64_t*, RebaseOffset, int> beg(g.nindex + r0, RebaseOffset{b});
        rocprim::transform_iterator<const int64_t*, RebaseOffset, int> end(g.nindex + r0 + 1, RebaseOffset{b});
        size_t need = 0;
        e = hipcub::DeviceSegmentedSort::SortPairs(nullptr, need, g.nlist + b, sval + b, pos_in, spos + b, items,
                                                    r1 - r0, beg, end, st);
        if (e != hipSuccess) break;
        if (need > tmp_bytes) {
            if (tmp) mf_free_device(tmp);
            tmp = nullptr;
            tmp_bytes = 0;
            e = hipMalloc(&tmp, need);
            if (e != hipSuccess) break;
            tmp_bytes = need;
        }
        e = hipcub::DeviceSegmentedSort::SortPairs(tmp, tmp_bytes, g.nlist + b, sval + b, pos_in, spos + b, items,
                                                    r1 - r0, beg, end, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (tmp) mf_free_device(tmp);
    mf_free_device(pos_in);
    if (e != hipSuccess) return e;
    k_check_row_index<<<S.blocks, 256, 0, st>>>(g.nindex, g.nlist, sval, spos, g.n, d_flags);
    return hipGetLastError();
}

// The sorted copy of compacted rows by merging (see k_merge_base): the
// deletion and tail lists of each row are sorted (they are a few entries a
// row), the old sorted order is merged with them entry by entry, and the
// result is checked.  S describes the graph before compaction (its rows, the
// deletions still marked, its sorted copy if it has one); nidx/nnl are the
// compacted rows.  Lists of more than 2^31 entries are refused.
namespace {
// begin (end: past = 1) of the i-th listed row's slice of a list, for the library sort
struct RowOffset {  // a listed row's list bounds; a slot naming no row (-1) is an empty list
    const int64_t* off;
    const int32_t* rows;
    int past;
    __host__ __device__ int operator()(const int& i) const { return rows[i] < 0 ? 0 : (int)off[rows[i] + past]; }
};

// Sorts the per-row lists of the listed rows: short ones a wave each, the
// rest with the library's segmented sort over just those rows.
hipError_t sort_lists(GraphState& S, const int64_t* off, const int32_t* rows, const u64* nrows, int64_t total,
                      const int32_t* kin, const int32_t* vin, int32_t* kout, int32_t* vout, const MergeLists& L) {
    hipStream_t st = S.stream;
    if (total == 0) return hipSuccess;
    // Without reading back how many lists are long: there are at most
    // total / (MF_SORT_WAVE + 1) of them (and at most one per row), and the
    // slots k_sort_small leaves unused name no row, so they sort nothing.
    const int64_t nbig = std::min<int64_t>(total / (MF_SORT_WAVE + 1), S.g.n);
    if (nbig) k_fill<int32_t><<<(unsigned)std::min<int64_t>((nbig + 255) / 256, 1024), 256, 0, st>>>(L.big, -1, nbig);
    k_fill<u64><<<1, 64, 0, st>>>(L.nbig, 0ull, 1);
    k_sort_small<<<S.blocks, 256, 0, st>>>(off, rows, nrows, kin, vin, kout, vout, L.big, L.nbig);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nbig == 0) return e;
    rocprim::counting_iterator<int> idx(0);
    rocprim::transform_iterator<rocprim::counting_iterator<int>, RowOffset, int> beg(idx, RowOffset{off, L.big, 0});
    rocprim::transform_iterator<rocprim::counting_iterator<int>, RowOffset, int> end(idx, RowOffset{off, L.big, 1});
    size_t need = 0;
    if (vin) e = hipcub::DeviceSegmentedSort::SortPairs(nullptr, need, kin, kout, vin, vout, (int)total, (int)nbig,
                                                         beg, end, st);
    else e = hipcub::DeviceSegmentedSort::SortKeys(nullptr, need, kin, kout, (int)total, (int)nbig, beg, end, st);
    if (e == hipSuccess && need > *L.tmp_bytes) {
        if (*L.tmp) mf_free_device(*L.tmp);
        *L.tmp = nullptr;
        *L.tmp_bytes = 0;
        e = hipMalloc(L.tmp, need);
        if (e == hipSuccess) *L.tmp_bytes = need;
    }
    if (e == hipSuccess) {
        size_t have = *L.tmp_bytes;
        if (vin) e = hipcub::DeviceSegmentedSort::SortPairs(*L.tmp, have, kin, kout, vin, vout, (int)total, (int)nbig,
                                                             beg, end, st);
        else e = hipcub::DeviceSegmentedSort::SortKeys(*L.tmp, have, kin, kout, (int)total, (int)nbig, beg, end, st);
    }
    return e;
}
}  // namespace

hipError_t mf_merge_row_index(GraphState& S, const MergeLists& L, int64_t ntail, const int64_t* nidx,
                              const int32_t* nnl, int32_t* nsval, int32_t* nspos, int* d_flags, bool check) {
    const DevGraph& g = S.g;
    hipStream_t st = S.stream;
    if (ntail > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    if (g.sval) {  // the sorted copy's marks (written by the copy, per merge-path chunk) made global
        mf_scan_excl(L.ccnt, L.ncst, L.cpre, L.cpart, st);
        k_words_fix_path<<<S.blocks, 256, 0, st>>>(L.tws, mf_tomb_words(g.m), L.cst, L.ncst, L.cpre);
    }
    hipError_t e = sort_lists(S, L.toff, L.trows, L.ntrows, ntail, L.tval, L.tidx, L.tsv, L.tsi, L);
    if (e == hipSuccess) {
        OldRows o{g.nindex, g.nlist, g.sval, g.spos};
        if (g.sval) k_merge_base<true><<<S.blocks, 256, 0, st>>>(o, g.n, L, L.tsv, L.tsi, ntail, nidx, nsval, nspos);
        else k_merge_base<false><<<S.blocks, 256, 0, st>>>(o, g.n, L, L.tsv, L.tsi, ntail, nidx, nsval, nspos);
        if (ntail) k_merge_tail_empty<<<S.blocks, 256, 0, st>>>(g.nindex, g.n, L, L.tsv, L.tsi, nidx, nsval, nspos);
        if (check) k_check_row_index<<<S.blocks, 256, 0, st>>>(nidx, nnl, nsval, nspos, g.n, d_flags);
        e = hipGetLastError();
    }
    return e;
}
