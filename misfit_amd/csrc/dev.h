// dev.h -- device-side data layout and helpers shared by every MISFIT kernel.
//
// HBM layout of one graph (everything stays device-resident for the run):
//   nindex[n+1]  int64   base CSR offsets (both .egr widths are widened once
//                         at load)
//   sval/spos    int32   when base rows are not ascending (input order, or
//                         after a compaction): every row sorted, with each
//                         entry's position in the row, for binary searches
//   nlist[m]     int32   base CSR entries.  A deleted base edge stays in
//                         place as its bitwise complement (~v < 0): a row read
//                         skips it with one sign test and re-inserting it
//                         restores the original position, which is what the
//                         reference's tombstone sets give
//                         (core/FlashGraph.h:175,196-233).
//   ov_*         per-vertex overlay run in a shared pool: inserted
//                         neighbours in insertion order; a removed insertion
//                         becomes ~v inside its run (the reference's
//                         added_adj vectors, core/FlashGraph.h:174,260-275).
//   ph_key/ph_val         open-addressed hash of inserted pairs
//                         (key lo<<32|hi; value 1 present, 0 removed).
//   mc[n]        uint32  bit 31 = MIS membership; bits 0..30 = number of MIS
//                         neighbours (self-loops excluded).  The count turns
//                         the reference's "any MIS member in this row" scans
//                         into one 4-byte gather.
//   per-batch epoch words (a word equal to the current batch epoch b means
//   "set in this batch", so nothing is ever cleared between batches):
//     seedb   = b   vertex is an endpoint of batch b (it has a 2-hop row)
//     wpotb   = b   vertex is max(u,v) of an insertion of batch b (may be
//                   evicted by the rule)
//     pub     = 4b+1 | 2    vertex may become uncovered during batch b
//                   (bit 1: it also has a row, i.e. it is admissible)
//     kw      = b<<32 | k   k = MIS neighbours of the vertex that may be
//                   evicted in batch b (counted for non-members next to
//                   members of W)
//     tagw    bitmap (not epoch words): the vertex has a count in kw for
//                   the current batch; the PU scan clears it as it reads it
//     evb     = b   vertex was evicted in batch b
//     nevb    = b   vertex is a neighbour of one evicted in batch b
//     pfb     = b   vertex is max(u,v) of an insertion of batch b whose other
//                   end may be evicted too (or whose edge is gone after the
//                   batch): only then can it be uncovered after an eviction
//     cstat   = 4b+s        completion state (1 open, 2 joined, 3 covered)
//     fa_*    per-batch index of admissible neighbours of long rows
//   mmark/cmark  u64 round<<32 | ~edge: earliest edge of the current
//                   speculative round that flips the vertex's membership
//                   (mmark) or changes its MIS-neighbour count (cmark).
//   cdec         u64 round<<32 | number of evictions next to the vertex
//                   speculated in the current round (an upper bound on how
//                   far earlier edges can lower its count).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MF_MISBIT 0x80000000u
#define MF_CMASK 0x7FFFFFFFu
#define MF_WAVE 64
#define MF_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

typedef unsigned long long u64;

struct DevGraph {
    int32_t n;
    int64_t m;
    const int64_t* nindex;
    int32_t* nlist;
    int64_t* ov_start;
    int32_t* ov_len;
    int32_t* ov_cap;
    int32_t* ov_alive;
    int32_t* ov_tomb;
    int32_t* pool;
    int64_t pool_cap;
    u64* pool_top;
    u64* ph_key;
    int32_t* ph_val;
    u64 ph_mask;
    uint32_t* mc;
    uint32_t* seedb;
    uint32_t* wpotb;
    uint32_t* pub;
    uint32_t* evb;
    u64* kw;
    uint32_t* tagw;  // bitmap of the vertices the k-count tagged (cleared by the PU scan)
    uint32_t* pfb;   // = b: an insertion of batch b may let this vertex become uncovered after its eviction
    uint32_t* nevb;  // = b: neighbour of a vertex evicted in batch b
    uint32_t* cstat;
    u64* mmark;
    u64* cmark;
    u64* cdec;       // round<<32 | evictions next to the vertex speculated this round
    int32_t* ne_head;
    uint32_t* fa_tag;  // = b: the row has admissible-index entries in batch b
    uint32_t* fa_n;    // entries counted for the index (reset by the allocation)
    int64_t* fa_start;
    int32_t* fa_fill;
    int sorted;      // every base row strictly ascending
    const int32_t* sval;  // rows not ascending: each row sorted (plain ids) ...
    const int32_t* spos;  // ... with each entry's position in its row
};

struct Row {
    const int32_t* base;
    int64_t blen;
    const int32_t* add;
    int64_t alen;
    __device__ __forceinline__ int64_t len() const { return blen + alen; }
    __device__ __forceinline__ int32_t at(int64_t j) const { return j < blen ? base[j] : add[j - blen]; }
};

// The four words of a row's bounds are loaded together: the run's start
// (every vertex has one, 0 until it gets a run) is read whatever the run's
// length, so no load waits for another; an empty run's pointer is never
// dereferenced.
__device__ __forceinline__ Row get_row(const DevGraph& g, int32_t x) {
    Row r;
    const int64_t s = g.nindex[x], e = g.nindex[x + 1];
    const int32_t al = g.ov_len[x];
    const int64_t os = g.ov_start[x];
    r.base = g.nlist + s;
    r.blen = e - s;
    r.alen = al;
    r.add = g.pool + os;
    return r;
}

__device__ __forceinline__ int64_t row_len(const DevGraph& g, int32_t x) {
    return (g.nindex[x + 1] - g.nindex[x]) + g.ov_len[x];
}

__device__ __forceinline__ int32_t decode(int32_t e) { return e < 0 ? ~e : e; }

__device__ __forceinline__ u64 pair_key(int32_t a, int32_t b) {
    uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((u64)lo << 32) | (u64)hi;
}

// SplitMix64 finaliser (Steele, Lea, Flood, OOPSLA 2014).
__device__ __forceinline__ u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Value of an inserted pair from probe slot s on: 1 present, 0 removed or
// never inserted.
__device__ __forceinline__ int32_t ph_lookup_from(const DevGraph& g, u64 key, u64 s) {
    while (true) {
        u64 k = g.ph_key[s];
        if (k == key) return g.ph_val[s];
        if (k == MF_EMPTY_KEY) return 0;
        s = (s + 1) & g.ph_mask;
    }
}

__device__ __forceinline__ int32_t ph_lookup(const DevGraph& g, u64 key) {
    return ph_lookup_from(g, key, mix64(key) & g.ph_mask);
}

// Slot of `key`, inserting it when absent.  The host keeps the load factor
// at or below one half, so the probe always ends.
__device__ __forceinline__ u64 ph_slot(const DevGraph& g, u64 key) {
    u64 s = mix64(key) & g.ph_mask;
    while (true) {
        u64 prev = atomicCAS(&g.ph_key[s], MF_EMPTY_KEY, key);
        if (prev == MF_EMPTY_KEY || prev == key) return s;
        s = (s + 1) & g.ph_mask;
    }
}

// Position of y (live or deleted) inside base row x, or -1.
__device__ __forceinline__ int64_t base_find(const DevGraph& g, int32_t x, int32_t y) {
    int64_t b = g.nindex[x], e = g.nindex[x + 1];
    if (g.sorted) {
        int64_t lo = b, hi = e;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            int32_t d = decode(g.nlist[mid]);
            if (d < y) lo = mid + 1;
            else hi = mid;
        }
        if (lo < e && decode(g.nlist[lo]) == y) return lo;
        return -1;
    }
    if (g.sval) {
        int64_t lo = b, hi = e;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (g.sval[mid] < y) lo = mid + 1;
            else hi = mid;
        }
        if (lo < e && g.sval[lo] == y) return b + g.spos[lo];
        return -1;
    }
    for (int64_t p = b; p < e; p++)
        if (decode(g.nlist[p]) == y) return p;
    return -1;
}

// base_find by a group of G lanes of one wave (lanes gbase .. gbase+G-1,
// `gl` = lane - gbase; the whole group calls it together and gets the same
// result): a (G+1)-ary search that probes G positions per step, so a hub row
// of 2^17 entries takes 6 dependent loads at G = 8 instead of 17.
template <int G>
__device__ __forceinline__ int64_t group_find(const DevGraph& g, int32_t x, int32_t y, int gl, int gbase) {
    const u64 gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    const int64_t b = g.nindex[x], e = g.nindex[x + 1];
    if (!g.sorted && !g.sval) {
        for (int64_t p0 = b; p0 < e; p0 += G) {
            const int64_t p = p0 + gl;
            const bool hit = p < e && decode(g.nlist[p]) == y;
            const u64 m = (__ballot(hit) >> gbase) & gmask;
            if (m) return p0 + __ffsll((long long)m) - 1;
        }
        return -1;
    }
    // the first index of the sorted view holding a value >= y lies in [lo, hi]
    // (hi = e: none)
    int64_t lo = b, hi = e;
    while (hi - lo >= G) {
        const int64_t span = hi - lo;
        const int64_t p = lo + span * (gl + 1) / (G + 1);  // lo <= p < hi
        const int32_t d = g.sorted ? decode(g.nlist[p]) : g.sval[p];
        const u64 m = (__ballot(d < y) >> gbase) & gmask;  // probes ascend: a prefix is below y
        const int cnt = __popcll(m);
        const int64_t nlo = cnt == 0 ? lo : lo + span * cnt / (G + 1) + 1;
        const int64_t nhi = cnt == G ? hi : lo + span * (cnt + 1) / (G + 1);
        lo = nlo;
        hi = nhi;
    }
    const int64_t p = lo + gl;  // at most G candidates: lo .. hi, below e
    const bool hit = p <= hi && p < e && (g.sorted ? decode(g.nlist[p]) : g.sval[p]) == y;
    const u64 m = (__ballot(hit) >> gbase) & gmask;
    if (!m) return -1;
    const int64_t q = lo + __ffsll((long long)m) - 1;
    return g.sorted ? q : b + g.spos[q];
}

// Position of y inside base row x and of x inside base row y (or -1): the
// two binary searches advance in lockstep so their dependent load chains
// overlap.
__device__ __forceinline__ void base_find_pair(const DevGraph& g, int32_t x, int32_t y, int64_t& px, int64_t& py) {
    if (!g.sorted) {
        px = base_find(g, x, y);
        py = base_find(g, y, x);
        return;
    }
    int64_t bx = g.nindex[x], ex = g.nindex[x + 1];
    int64_t by = g.nindex[y], ey = g.nindex[y + 1];
    int64_t lx = bx, hx = ex, ly = by, hy = ey;
    while (lx < hx || ly < hy) {
        int64_t mx = (lx + hx) >> 1, my = (ly + hy) >> 1;
        int32_t dx = lx < hx ? decode(g.nlist[mx]) : 0;
        int32_t dy = ly < hy ? decode(g.nlist[my]) : 0;
        if (lx < hx) {
            if (dx < y) lx = mx + 1;
            else hx = mx;
        }
        if (ly < hy) {
            if (dy < x) ly = my + 1;
            else hy = my;
        }
    }
    px = (lx < ex && decode(g.nlist[lx]) == y) ? lx : -1;
    py = (ly < ey && decode(g.nlist[ly]) == x) ? ly : -1;
}

// Adjacency in the current graph (base minus deletions plus insertions): the
// device form of HybridGraph::isAdjacent (core/FlashGraph.h:291-301).  The
// base part searches the shorter of the two rows (rows are symmetric).
__device__ __forceinline__ bool adjacent(const DevGraph& g, int32_t x, int32_t y) {
    if (ph_lookup(g, pair_key(x, y)) == 1) return true;
    if ((g.sorted || g.sval) && (g.nindex[y + 1] - g.nindex[y]) < (g.nindex[x + 1] - g.nindex[x])) {
        int32_t t = x;
        x = y;
        y = t;
    }
    int64_t p = base_find(g, x, y);
    return p >= 0 && g.nlist[p] >= 0;
}

// Flat position of live neighbour z inside row(x) (base part first, then
// the overlay run), or -1.
__device__ __forceinline__ int64_t row_pos(const DevGraph& g, int32_t x, int32_t z) {
    int64_t p = base_find(g, x, z);
    if (p >= 0 && g.nlist[p] >= 0) return p - g.nindex[x];
    int32_t al = g.ov_len[x];
    if (al) {
        const int32_t* a = g.pool + g.ov_start[x];
        for (int32_t i = 0; i < al; i++)
            if (a[i] == z) return (g.nindex[x + 1] - g.nindex[x]) + i;
    }
    return -1;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (MF_WAVE - 1); }

__device__ __forceinline__ int first_lane(u64 mask) { return mask ? __ffsll((long long)mask) - 1 : -1; }

// Appends v for every active lane with pred set, with one atomic per wave.
// Call from wave-uniform or divergent code alike: only active lanes take part.
__device__ __forceinline__ void wave_push(int32_t* list, u64* ctr, int32_t v, bool pred) {
    u64 bal = __ballot(pred);
    if (!bal) return;
    int lane = lane_id();
    int leader = first_lane(bal);
    u64 base = 0;
    if (lane == leader) base = atomicAdd(ctr, (u64)__popcll(bal));
    base = (u64)__shfl((long long)base, leader, MF_WAVE);
    if (pred) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = v;
}

// wave_push of up to two values per lane (v1 when p1, v2 when p2), with one
// atomic per wave.
__device__ __forceinline__ void wave_push2(int32_t* list, u64* ctr, int32_t v1, bool p1, int32_t v2, bool p2) {
    const u64 b1 = __ballot(p1), b2 = __ballot(p2);
    if (!(b1 | b2)) return;
    const int lane = lane_id();
    const u64 below = (1ull << lane) - 1ull;
    const int leader = first_lane(__ballot(1));
    u64 base = 0;
    if (lane == leader) base = atomicAdd(ctr, (u64)(__popcll(b1) + __popcll(b2)));
    base = (u64)__shfl((long long)base, leader, MF_WAVE);
    const u64 at = base + (u64)(__popcll(b1 & below) + __popcll(b2 & below));
    if (p1) list[at] = v1;
    if (p2) list[at + (p1 ? 1 : 0)] = v2;
}

// Wave-wide sum; every lane of the wave must be active.
__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
    for (int d = MF_WAVE / 2; d > 0; d >>= 1) v += (u64)__shfl_xor((long long)v, d, MF_WAVE);
    return v;
}

// Adds v to *ctr with one atomic per wave; every lane must be active.
__device__ __forceinline__ void wave_add(u64* ctr, u64 v) {
    u64 t = wave_sum_u64(v);
    if (t && lane_id() == 0) atomicAdd(ctr, t);
}

// Wave-wide inclusive prefix sum.
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    int l = lane_id();
#pragma unroll
    for (int d = 1; d < MF_WAVE; d <<= 1) {
        int64_t t = __shfl_up(v, d, MF_WAVE);
        if (l >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ u64 wave_min_u64(u64 v) {
#pragma unroll
    for (int d = MF_WAVE / 2; d > 0; d >>= 1) {
        u64 o = __shfl_xor(v, d, MF_WAVE);
        v = o < v ? o : v;
    }
    return v;
}

// Grid coordinates for grid-stride phases (one phase body serves both the
// grid-wide kernels and the single-workgroup finisher).
struct GC {
    int64_t tid, nthreads, wave, nwaves;
    int lane;
};

// Every grid-stride kernel runs workgroups of MF_BLOCK threads.  The size is
// a constant here, not blockDim.x: reading the workgroup size is a load from
// the dispatch's hidden arguments that every wave would wait for first.
#define MF_BLOCK 256
__device__ __forceinline__ GC grid_coords() {
    GC c;
    c.tid = (int64_t)blockIdx.x * MF_BLOCK + threadIdx.x;
    c.nthreads = (int64_t)gridDim.x * MF_BLOCK;
    c.wave = c.tid / MF_WAVE;
    c.nwaves = c.nthreads / MF_WAVE;
    c.lane = lane_id();
    return c;
}
