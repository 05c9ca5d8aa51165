// k_batch.hip -- one update batch on the device: overlay apply, the set of
// vertices the batch may uncover, the exact sequential repair rule evaluated
// in speculative rounds, and the maximality completion.
//
// Semantics (reference = anonymous-artifacts/MISFIT):
//   * the whole batch is applied to the overlay first
//     (src/misfit/misfit.cu:704-724), so the repair reads the post-batch
//     graph, which is static while membership evolves;
//   * the rule is applied per edge in batch order
//     (src/misfit/misfit.cu:283-347):
//       insertion, both ends in the MIS: evict w = max(u,v), then admit the
//         first vertex x of row(w) (BFS order: [w] ++ N1(w) ++ N2(w)) that is
//         outside the MIS, has a row, and whose row holds no MIS member;
//       deletion, exactly one end a in the MIS: admit the other end b when
//         row(b) holds no MIS member other than a;
//     where row(s) = [s] ++ N1 ++ N2 for batch endpoints and row(x) = adj(x)
//     for other neighbours of endpoints (src/misfit/misfit.cu:158-202,
//     550-558); nobody else has a row;
//   * completion: over C = endpoints ∪ evicted ∪ N(evicted), in ascending
//     id, a vertex joins when it is outside the MIS with no MIS neighbour.
//
// How the GPU evaluates the sequential rule exactly:
//   1. PU ("potentially uncovered") is a superset of the vertices that are at
//      some moment of the batch outside the MIS with no MIS neighbour.  Only
//      evictions remove members, and only W = {max(u,v) of insertions} can be
//      evicted.  So, with c(z) = MIS neighbours and k(z) = MIS neighbours in W
//      at batch start:
//        - a non-member z with c(z) > k(z) keeps a member neighbour forever;
//        - a member z outside W stays a member;
//        - a member z in W becomes a non-member only when some insertion
//          (z, o), z > o, evicts it, and o, a member at that moment and a
//          neighbour of z after the batch, covers z until o itself is
//          evicted, which needs o in W.
//      PU = {z not in MIS : c(z) = k(z)}
//         ∪ {z in W ∩ MIS : some insertion (z, o), z > o, has o in W or the
//            pair absent after the batch}.
//      Only PU members with a row (ADM) can ever be admitted, and a vertex
//      outside PU always blocks a 2-hop test, so the repair reads nothing
//      else.
//   2. BFS discovery order of row(w) equals first-occurrence order of
//      [w] ++ adj(w) ++ concat_{x1 in adj(w)} adj(x1), so the first admissible
//      vertex is found by scanning that flat sequence; long adj(x1) rows are
//      replaced by a per-batch index of their ADM members with positions.
//   3. Edges are evaluated speculatively in rounds: every pending edge of a
//      window is evaluated against the current state; the writes it would do
//      are marked per vertex with the earliest writing edge; a second walk
//      finds the first edge that read a vertex written by an earlier edge of
//      the round.  Every edge before it saw exactly the sequential state and
//      commits; the rest are re-evaluated next round (the first pending edge
//      never conflicts, so every round makes progress).
#include <mutex>

#include "batch.h"

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ bool is_adm(const BP& P, int32_t x) { return P.g.pub[x] == 4u * P.b + 3u; }

// z joins PU when pred holds (first time in this batch: appended to the PU
// list).  Every list below holds distinct vertices, so n entries suffice.
__device__ __forceinline__ void pu_add(const BP& P, int32_t z, bool pred) {
    bool first = false;
    if (pred) {
        uint32_t want = 4u * P.b + 1u;
        first = atomicMax(&P.g.pub[z], want) < want;
    }
    wave_push(P.pu, &P.ctr[MF_CS * (C_PU)], z, first);
}

__device__ __forceinline__ bool mk_before(const u64* mk, int32_t x, uint32_t R, int32_t e) {
    u64 m = mk[x];
    return (uint32_t)(m >> 32) == R && (int64_t)(0xFFFFFFFFu - (uint32_t)m) < (int64_t)e;
}

// Validation of one read of mc[x] that returned `val` by edge e.  The
// decision taken from it holds while the membership bit is unchanged and the
// count stays at or above T (T = 0: any change of the count matters;
// MF_T_MIS: the count does not matter).  With MF_T_NOFLIP (on a count
// threshold of a non-member) a flip of x does not matter either: a
// non-member can only join the MIS, and a member blocks wherever a
// non-member with that many MIS neighbours blocks.  The read conflicts when
// an earlier edge of the round flipped the bit, or changed the count and the
// evictions speculated next to x this round could take it below T.
#define MF_T_MIS 0xFFFFFFFFu
#define MF_T_NOFLIP 0x80000000u
__device__ __forceinline__ void vread(const BP& P, int32_t x, uint32_t val, uint32_t T, uint32_t R, int32_t e,
                                      bool& conf) {
    const bool noflip = T != MF_T_MIS && (T & MF_T_NOFLIP);
    if (noflip) T &= ~MF_T_NOFLIP;
    if (!noflip && mk_before(P.g.mmark, x, R, e)) {
        conf = true;
        return;
    }
    if (T == MF_T_MIS || !mk_before(P.g.cmark, x, R, e)) return;
    uint32_t cnt = val & MF_CMASK;
    if (T == 0 || cnt < T) {
        conf = true;
        return;
    }
    u64 d = P.g.cdec[x];
    uint32_t dec = ((uint32_t)(d >> 32) == R) ? (uint32_t)d : 0u;
    if ((u64)cnt < (u64)T + dec) conf = true;
}

// Per-wave log of the guarded reads of a speculation round: a fixed segment
// of MF_RLOG 16-byte slots per wave, so logging needs no global atomics; the
// cursor lives in LDS.  Slot 0 is the header (record count in .x), written
// when the wave's speculation ends; slots 1 .. MF_RLOG-1 hold the records
// (vertex, value read, threshold, edge), so the header and the first seven
// records share one 128-byte line.  A read that does not fit marks its edge
// for a full re-walk instead.
struct ReadLog {
    uint4* rec;        // this wave's segment
    int* cur;          // this wave's cursor (LDS)
    uint32_t* relog;   // per edge: = round when its log is incomplete
};

__device__ __forceinline__ void vlog(const ReadLog& lg, int32_t x, uint32_t val, uint32_t T, int32_t e, uint32_t R) {
    u64 bal = __ballot(1);
    const int lane = lane_id();
    const int leader = first_lane(bal);
    int base = 0;
    if (lane == leader) base = atomicAdd(lg.cur, __popcll(bal));
    base = __shfl(base, leader, MF_WAVE);
    int pos = base + __popcll(bal & ((1ull << lane) - 1ull));
    if (pos < MF_RLOG - 1) lg.rec[1 + pos] = make_uint4((uint32_t)x, val, T, (uint32_t)e);
    else lg.relog[e] = R;
}

// A guarded read: vertex, the mc word read, threshold (see vread).  x < 0:
// none (the decision rests on per-batch static data only).
struct Rd {
    int32_t x;
    uint32_t val, T;
};
__device__ __forceinline__ Rd rd(int32_t x, uint32_t val, uint32_t T) {
    Rd r;
    r.x = x;
    r.val = val;
    r.T = T;
    return r;
}
__device__ __forceinline__ Rd rd_none() { return rd(-1, 0u, 0u); }

// Where the reads an outcome rests on go: nowhere (plain speculation),
// checked against this round's marks (validation by re-walk), or appended to
// the wave's read log (speculation before a flat validation).
//
// Log records are held in registers (up to four per lane) until
// flush_records: a vector memory counter counts stores with loads, so a
// store issued before a load holds up the load's use until the store is
// done, and an evaluation's records would otherwise delay its own later
// loads.
#define REC_NONE 0
#define REC_CHECK 1
#define REC_LOG 2
struct Recorder {
    int mode;
    ReadLog lg;
    uint32_t R;
    int32_t e;
    bool conf;  // REC_CHECK: a recorded read conflicts (per lane)
    int np = 0; // REC_LOG: records held back
    Rd p0, p1, p2, p3;
};

__device__ __forceinline__ void record(const BP& P, Recorder& rc, const Rd& r) {
    if (r.x < 0) return;
    if (rc.mode == REC_CHECK) {
        vread(P, r.x, r.val, r.T, rc.R, rc.e, rc.conf);
    } else if (rc.mode == REC_LOG) {
        // named slots, not an array: a lane-dependent index would live in scratch memory
        if (rc.np == 0) rc.p0 = r;
        else if (rc.np == 1) rc.p1 = r;
        else if (rc.np == 2) rc.p2 = r;
        else if (rc.np == 3) rc.p3 = r;
        else vlog(rc.lg, r.x, r.val, r.T, rc.e, rc.R);
        if (rc.np < 4) rc.np++;
    }
}

// Appends the records held back to the wave's log (any subset of lanes may
// hold some).
__device__ __forceinline__ void flush_records(Recorder& rc) {
    if (rc.mode != REC_LOG) return;
    if (rc.np > 0) vlog(rc.lg, rc.p0.x, rc.p0.val, rc.p0.T, rc.e, rc.R);
    if (rc.np > 1) vlog(rc.lg, rc.p1.x, rc.p1.val, rc.p1.T, rc.e, rc.R);
    if (rc.np > 2) vlog(rc.lg, rc.p2.x, rc.p2.val, rc.p2.T, rc.e, rc.R);
    if (rc.np > 3) vlog(rc.lg, rc.p3.x, rc.p3.val, rc.p3.T, rc.e, rc.R);
    rc.np = 0;
}

// Round-tagged counter word R<<32 | count.  Rounds only grow, so raising
// the word to R<<32 and then adding one counts this round's additions: the
// two no-return atomics of one lane reach the word in program order.
__device__ __forceinline__ void round_count_add(u64* w, uint32_t R) {
    atomicMax(w, (u64)R << 32);
    atomicAdd(w, 1ull);
}

// ---------------------------------------------------------------- prologue
// P1b: the pair's adjacency before the batch (base positions, overlay
// state), for P2.  Nothing of the graph changes before P2, so these
// dependent searches run in waves of their own beside the claims instead of
// after them: 16 lanes per update, 8 searching each row (group_find).
#define MF_PAIR_LANES 16
__device__ __forceinline__ void search_pair(const BP& P, int64_t i, int gl) {
    const DevGraph& g = P.g;
    const int32_t a = P.bu[i], bb = P.bv[i];
    const int32_t lo = a < bb ? a : bb, hi = a < bb ? bb : a;
    const bool first = gl < MF_PAIR_LANES / 2;
    const int gbase = lane_id() & ~(MF_PAIR_LANES / 2 - 1);
    // the overlay hash's first probe, loaded before the searches and used
    // after them when the pair is not in the base
    const u64 key = pair_key(lo, hi);
    const u64 h0 = mix64(key) & g.ph_mask;
    u64 k0 = 0;
    int32_t v0 = 0;
    if (gl == 0) {
        k0 = g.ph_key[h0];
        v0 = g.ph_val[h0];
    }
    const int64_t p = group_find<MF_PAIR_LANES / 2>(g, first ? lo : hi, first ? hi : lo, gl & (MF_PAIR_LANES / 2 - 1), gbase);
    const int64_t phi = __shfl((long long)p, lane_id() + MF_PAIR_LANES / 2, MF_WAVE);
    if (gl != 0) return;
    const int64_t plo = p;
    uint8_t fl = 0;
    if (plo >= 0) fl = g.nlist[plo] < 0 ? 2 : 0;
    else if (k0 == key) fl = v0 == 1 ? 1 : 0;
    else if (k0 != MF_EMPTY_KEY) fl = ph_lookup_from(g, key, (h0 + 1) & g.ph_mask) == 1 ? 1 : 0;
    P.bt_plo[i] = plo;
    P.bt_phi[i] = phi;
    P.bt_fl[i] = fl;
}

// Threads of the claim kernel: [0, U) claim, the searches start at the next
// wave boundary (MF_PAIR_LANES per update).
__host__ __device__ __forceinline__ int64_t claim_search_base(int64_t U) { return (U + MF_WAVE - 1) / MF_WAVE * MF_WAVE; }
__host__ __device__ __forceinline__ int64_t claim_threads(int64_t U) { return claim_search_base(U) + MF_PAIR_LANES * U; }

// P1: claim pairs, tag endpoints and possible evictees, and search every
// update's pair.  A search group's lanes share one wave and one loop
// iteration: thread t is lane t mod 64 of its wave (blocks are whole waves).
__device__ void ph_claim(const BP& P, const GC& c) {
    const int64_t s0 = claim_search_base(P.U), end = claim_threads(P.U);
    for (int64_t t = c.tid; t < end; t += c.nthreads) {
        if (t >= s0) {
            search_pair(P, (t - s0) / MF_PAIR_LANES, (int)((t - s0) % MF_PAIR_LANES));
            continue;
        }
        if (t >= P.U) continue;
        const int64_t i = t;
        int32_t a = P.bu[i], bb = P.bv[i];
        const bool ins = P.bins[i] != 0;
        const int32_t w = a > bb ? a : bb;
        // the tag exchanges are independent of the pair table: all in flight
        // together with its insertion
        const bool fa = atomicExch(&P.g.seedb[a], P.b) != P.b;
        const bool fb = atomicExch(&P.g.seedb[bb], P.b) != P.b;
        const bool fw = ins && atomicExch(&P.g.wpotb[w], P.b) != P.b;
        u64 key = pair_key(a, bb);
        u64 s = mix64(key) & P.bt_mask;
        while (true) {
            u64 prev = atomicCAS(&P.bt_key[s], MF_EMPTY_KEY, key);
            if (prev == MF_EMPTY_KEY || prev == key) break;
            s = (s + 1) & P.bt_mask;
        }
        int32_t prevh = atomicExch(&P.bt_head[s], (int32_t)i);
        P.bt_nxt[i] = prevh;
        P.bt_slot[i] = (int32_t)s;
        wave_push2(P.seeds, &P.ctr[MF_CS * (C_SEEDS)], a, fa, bb, fb);
        wave_push(P.wpot, &P.ctr[MF_CS * (C_WPOT)], w, fw);
    }
}

__device__ __forceinline__ void set_base_deleted(const DevGraph& g, int32_t x, int64_t p, bool del) {
    int32_t e = g.nlist[p];
    if (del && e >= 0) {
        g.nlist[p] = ~e;
        atomicAdd(&g.ov_tomb[x], 1);
    } else if (!del && e < 0) {
        g.nlist[p] = ~e;
        atomicSub(&g.ov_tomb[x], 1);
    }
}

__device__ __forceinline__ void kill_run_entry(const DevGraph& g, int32_t x, int32_t y) {
    int32_t len = g.ov_len[x];
    int32_t* r = g.pool + g.ov_start[x];
    for (int32_t i = 0; i < len; i++) {
        if (r[i] == y) {
            r[i] = ~y;
            atomicSub(&g.ov_alive[x], 1);
            return;
        }
    }
}

// P2: one thread per distinct pair replays the pair's updates in batch order
// with the overlay's rules (core/FlashGraph.h:196-212 via misfit.cu:708-714)
// and applies the net effect.
__device__ void ph_resolve(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    u64 my_ops = 0;
    for (int64_t i = c.tid; i < P.U; i += c.nthreads) {
        // new overlay entries of the pair headed by update i go to 2i, 2i+1
        P.ne_x[2 * i] = -1;
        P.ne_x[2 * i + 1] = -1;
        // everything update i's own index gives is loaded beside its slot,
        // before the head test (and the endpoints' count words with it)
        const int32_t slot = P.bt_slot[i];
        const int32_t a = P.bu[i], bb = P.bv[i];
        const int64_t plo = P.bt_plo[i], phi = P.bt_phi[i];
        const uint8_t fl = P.bt_fl[i];
        const uint32_t ma = g.mc[a], mb = g.mc[bb];
        if (P.bt_head[slot] != (int32_t)i) continue;
        int32_t lo = a < bb ? a : bb, hi = a < bb ? bb : a;
        bool in_base = plo >= 0;
        bool tomb0 = (fl & 2) != 0;
        u64 key = pair_key(lo, hi);
        bool alive0 = (fl & 1) != 0;
        bool adj0 = (in_base && !tomb0) || alive0;
        bool tomb = tomb0, alive = alive0, adj = adj0, old_killed = false;
        int32_t last_ins = -1;
        int64_t ops = 0;
        int64_t prev_idx = -1;
        while (true) {
            int64_t cur = INT64_MAX;
            for (int32_t j = (int32_t)i; j != -1; j = P.bt_nxt[j])
                if (j > prev_idx && j < cur) cur = j;
            if (cur == INT64_MAX) break;
            prev_idx = cur;
            bool ins = P.bins[cur] != 0;
            if (ins && !adj) {
                if (in_base && tomb) tomb = false;
                else { alive = true; last_ins = (int32_t)cur; }
                adj = true;
                ops++;
            } else if (!ins && adj) {
                if (alive) {
                    alive = false;
                    if (last_ins < 0) old_killed = true;
                    last_ins = -1;
                } else {
                    tomb = true;
                }
                adj = false;
                ops++;
            }
        }
        P.bt_key[slot] = MF_EMPTY_KEY;
        P.bt_head[slot] = -1;
        my_ops += (u64)ops;
        if (adj)  // the pair's adjacency after the batch, for P5a
            for (int32_t j = (int32_t)i; j != -1; j = P.bt_nxt[j]) P.bt_fl[j] |= 4;
        if (in_base && tomb != tomb0) {
            set_base_deleted(g, lo, plo, tomb);
            if (phi >= 0) set_base_deleted(g, hi, phi, tomb);
        }
        if (old_killed) {
            kill_run_entry(g, lo, hi);
            kill_run_entry(g, hi, lo);
        }
        // new entries, chained per vertex: the head exchanges are issued
        // before the hash slot's claim (independent), their results stored
        // after it (nothing in this kernel reads the chains)
        const bool fresh = alive && last_ins >= 0;
        const int64_t k = 2 * i;
        int32_t nlo = -1, nhi = -1;
        if (fresh) {
            nlo = atomicExch(&g.ne_head[lo], (int32_t)k);
            nhi = atomicExch(&g.ne_head[hi], (int32_t)(k + 1));
        }
        if (!in_base && (alive != alive0 || last_ins >= 0)) {
            u64 s = ph_slot(g, key);
            g.ph_val[s] = alive ? 1 : 0;
        }
        if (fresh) {
            P.ne_x[k] = lo; P.ne_y[k] = hi; P.ne_seq[k] = last_ins;
            P.ne_x[k + 1] = hi; P.ne_y[k + 1] = lo; P.ne_seq[k + 1] = last_ins;
            P.ne_nxt[k] = nlo;
            P.ne_nxt[k + 1] = nhi;
        }
        if (adj != adj0) {
            uint32_t d = adj ? 1u : 0xFFFFFFFFu;  // +1 or -1 on the count bits
            const uint32_t mlo = a < bb ? ma : mb, mhi = a < bb ? mb : ma;
            if (mhi & MF_MISBIT) atomicAdd(&g.mc[lo], d);
            if (mlo & MF_MISBIT) atomicAdd(&g.mc[hi], d);
        }
    }
    wave_add(&P.ctr[MF_CS * (C_DELTA_OPS)], my_ops);
}

// P4: the chain owner appends the vertex's new entries to its run in batch
// order, growing the run by doubling inside the pool.
__device__ void ph_append(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    int64_t ne = 2 * P.U;
    // whole waves step together so the pool reservation below is one atomic
    // per wave (most appending vertices need a new run, and they all
    // reserve from one counter)
    for (int64_t k0 = c.tid - c.lane; k0 < ne; k0 += c.nthreads) {
        const int64_t k = k0 + c.lane;
        // everything that depends on k alone first, then everything that
        // depends on x: a vertex with one new entry (the usual case) needs
        // no chain walk, so its append is two loads and the pool
        // reservation deep
        int32_t x = -1, nxt = -1, y0 = 0;
        if (k < ne) {
            x = P.ne_x[k];
            nxt = P.ne_nxt[k];
            y0 = P.ne_y[k];
        }
        int32_t hd = -1, len = 0, cap = 0;
        int64_t ost = 0;
        if (x >= 0) {
            hd = g.ne_head[x];
            len = g.ov_len[x];
            cap = g.ov_cap[x];
            ost = g.ov_start[x];
        }
        const bool head = x >= 0 && hd == (int32_t)k;
        int32_t cnt = 0, nc = 0;
        if (head) {
            cnt = 1;
            for (int32_t j = nxt; j != -1; j = P.ne_nxt[j]) cnt++;
            if (len + cnt > cap) {
                nc = 2;
                while (nc < len + cnt) nc <<= 1;
            }
        }
        const int64_t incl = wave_incl_scan(nc);
        const int64_t total = __shfl(incl, MF_WAVE - 1, MF_WAVE);
        u64 base = 0;
        if (c.lane == 0 && total) base = atomicAdd(g.pool_top, (u64)total);
        base = (u64)__shfl((long long)base, 0, MF_WAVE);
        if (!head) continue;
        if (nc) {
            const u64 st = base + (u64)(incl - nc);
            if ((int64_t)(st + nc) > g.pool_cap) {
                atomicExch(&P.ctr[MF_CS * (C_POOL_OVF)], 1ull);
                g.ne_head[x] = -1;
                continue;
            }
            if (len) {
                const int32_t* src = g.pool + ost;
                for (int32_t t = 0; t < len; t++) g.pool[st + t] = src[t];
            }
            ost = (int64_t)st;
            g.ov_start[x] = ost;
            g.ov_cap[x] = nc;
        }
        int32_t* run = g.pool + ost;
        if (cnt == 1) {
            run[len] = y0;
        } else {
            int32_t prev_seq = -1;
            for (int32_t t = 0; t < cnt; t++) {
                int32_t best = -1, bseq = INT32_MAX;
                for (int32_t j = (int32_t)k; j != -1; j = P.ne_nxt[j]) {
                    int32_t sq = P.ne_seq[j];
                    if (sq > prev_seq && sq < bseq) { bseq = sq; best = j; }
                }
                run[len + t] = P.ne_y[best];
                prev_seq = bseq;
            }
        }
        g.ov_len[x] = len + cnt;
        atomicAdd(&g.ov_alive[x], cnt);
        g.ne_head[x] = -1;
    }
}

// P5a: flag members of W that could become uncovered after an eviction, and
// cut the rows of W's MIS members into chunks for the counting pass.
__device__ void ph_partner(const BP& P, const GC& c) {
    for (int64_t i = c.tid; i < P.U; i += c.nthreads) {
        const bool ins = P.bins[i] != 0;
        const int32_t a = P.bu[i], bb = P.bv[i];
        const uint8_t fl = P.bt_fl[i];
        if (!ins) continue;
        int32_t w = a > bb ? a : bb, o = a > bb ? bb : a;
        // (the store is idempotent: no test of pfb[w] first)
        if (P.g.wpotb[o] == P.b || !(fl & 4)) P.g.pfb[w] = P.b;  // bit 2: adjacent after the batch (P2)
    }
    int64_t nw = (int64_t)P.ctr[MF_CS * (C_WPOT)];
    // whole waves step together so the reservation below is wave-wide
    for (int64_t i0 = c.tid - c.lane; i0 < nw; i0 += c.nthreads) {
        int64_t i = i0 + c.lane;
        int32_t w = i < nw ? P.wpot[i] : -1;
        int64_t nch = 0;
        if (w >= 0 && (P.g.mc[w] & MF_MISBIT)) nch = (row_len(P.g, w) + MF_KCH - 1) / MF_KCH;
        int64_t incl = wave_incl_scan(nch);
        int64_t total = __shfl(incl, MF_WAVE - 1, MF_WAVE);
        u64 base = 0;
        if (c.lane == 0 && total) base = atomicAdd(&P.ctr[MF_CS * (C_KTASK)], (u64)total);
        base = (u64)__shfl((long long)base, 0, MF_WAVE);
        u64 t = base + (u64)(incl - nch);
        if (nch && (int64_t)(t + nch) > P.ktask_cap) {
            atomicExch(&P.ctr[MF_CS * (C_GUARD)], 3ull);
            continue;
        }
        for (int64_t k = 0; k < nch; k++) P.ktask[t + k] = ((u64)(uint32_t)w << 32) | (u64)k;
    }
}

// P5: k(z), the MIS neighbours of z that may be evicted, for non-members z
// next to members of W∩MIS (one wave per 64-entry chunk of such a row).  The
// count word is raised to the batch epoch and incremented with two no-return
// atomics (neighbours are mostly distinct, and nothing waits on the result);
// a no-return atomic OR tags z in the scan's bitmap.  Flagged members of W, and
// endpoints whose deleted edge took their last MIS neighbour, join PU here.
__device__ void ph_kcount(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    int64_t nw = (int64_t)P.ctr[MF_CS * (C_WPOT)];
    for (int64_t i = c.tid; i < nw; i += c.nthreads) {
        int32_t w = P.wpot[i];
        pu_add(P, w, (g.mc[w] & MF_MISBIT) && g.pfb[w] == P.b);
    }
    int64_t ns = (int64_t)P.ctr[MF_CS * (C_SEEDS)];
    for (int64_t i = c.tid; i < ns; i += c.nthreads) {
        int32_t s = P.seeds[i];
        pu_add(P, s, g.mc[s] == 0);
    }
    int64_t nt = (int64_t)P.ctr[MF_CS * (C_KTASK)];
    if (nt > P.ktask_cap) nt = P.ktask_cap;
    const u64 ep = (u64)P.b << 32;
    for (int64_t i = c.wave; i < nt; i += c.nwaves) {
        u64 task = P.ktask[i];
        int32_t w = (int32_t)(task >> 32);
        int64_t j = (int64_t)(uint32_t)task * MF_KCH + c.lane;
        Row r = get_row(g, w);
        if (j >= r.len()) continue;
        int32_t z = r.at(j);
        if (z < 0 || z == w) continue;
        uint32_t mz = g.mc[z];
        // c(z) <= |W| is necessary for all MIS neighbours to be in W
        if ((mz & MF_MISBIT) || (u64)(mz & MF_CMASK) > (u64)nw) continue;
        atomicMax(&g.kw[z], ep);
        atomicAdd(&g.kw[z], 1ull);
        atomicOr(&g.tagw[z >> 5], 1u << (z & 31));
    }
}

// P6: PU: a non-member with no MIS neighbour, or a tagged one all of whose
// MIS neighbours may be evicted (c(z) = k(z)).  The tags are a bitmap that
// this pass clears as it reads it, so it is zero at every batch start.
__device__ __forceinline__ void pu_scan_one(const BP& P, int32_t x, uint32_t m, bool tagged) {
    bool p = m == 0;
    if (!p && tagged && !(m & MF_MISBIT)) p = (m & MF_CMASK) == (uint32_t)(P.g.kw[x] & 0xFFFFFFFFull);
    pu_add(P, x, p);
}

// A tagged vertex: its count word and k-count loaded together.
__device__ __forceinline__ void pu_scan_tagged(const BP& P, int32_t x) {
    const uint32_t m = P.g.mc[x];
    const uint32_t k = (uint32_t)(P.g.kw[x] & 0xFFFFFFFFull);
    pu_add(P, x, m == 0 || (!(m & MF_MISBIT) && (m & MF_CMASK) == k));
}

// With a maximal MIS before the batch (P.pu_list), a vertex outside the MIS
// with no MIS neighbour after the overlay update is an endpoint of a deleted
// edge (the seeds, added in P5), so only the tagged vertices are tested: a
// thread per bitmap word.  Otherwise every vertex, four per thread; the
// eight lanes that share a bitmap word sit in one wave, which reads the word
// before its first lane clears it.
__device__ void ph_pu_scan(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    const int64_t nwords = ((int64_t)g.n + 31) / 32;
    if (P.pu_list) {
        for (int64_t q = c.tid; q < nwords; q += c.nthreads) {
            uint32_t w = g.tagw[q];
            if (!w) continue;
            g.tagw[q] = 0;
            while (w) {
                const int32_t x = (int32_t)(32 * q + __ffs(w) - 1);
                w &= w - 1;
                pu_scan_tagged(P, x);
            }
        }
        return;
    }
    const int64_t full = (int64_t)g.n / 32;  // whole words
    const uint4* mc4 = reinterpret_cast<const uint4*>(g.mc);
    for (int64_t q = c.tid; q < 8 * full; q += c.nthreads) {
        const uint4 m = mc4[q];
        const uint32_t w = g.tagw[q >> 3];
        const uint32_t t = w >> (4 * (q & 7));
        const int32_t x = (int32_t)(4 * q);
        pu_scan_one(P, x, m.x, t & 1u);
        pu_scan_one(P, x + 1, m.y, t & 2u);
        pu_scan_one(P, x + 2, m.z, t & 4u);
        pu_scan_one(P, x + 3, m.w, t & 8u);
        if ((q & 7) == 0 && w) g.tagw[q >> 3] = 0;
    }
    if (c.tid == 0 && full < nwords) {  // the last, partial word
        const uint32_t w = g.tagw[full];
        for (int64_t x = 32 * full; x < g.n; x++) pu_scan_one(P, (int32_t)x, g.mc[x], (w >> (x & 31)) & 1u);
        if (w) g.tagw[full] = 0;
    }
}

// P7: a PU member is admissible (ADM) when it has a row: it is an endpoint
// or next to one.  P8: for each ADM member, count it in every adjacency row
// it sits in (the rows the eviction scan reads through the index).
__device__ void ph_fa_count(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    int64_t np = (int64_t)P.ctr[MF_CS * (C_PU)];
    if (np > P.pu_cap) np = P.pu_cap;
    for (int64_t i = c.wave; i < np; i += c.nwaves) {
        int32_t z = P.pu[i];
        Row r = get_row(g, z);
        int64_t L = r.len();
        bool has = g.seedb[z] == P.b;
        for (int64_t j0 = 0; j0 < L && !has; j0 += MF_WAVE) {
            int64_t j = j0 + c.lane;
            bool s = false;
            if (j < L) {
                int32_t y = r.at(j);
                s = y >= 0 && g.seedb[y] == P.b;
            }
            has = __ballot(s) != 0;
        }
        if (!has) continue;
        if (c.lane == 0) {
            g.pub[z] = 4u * P.b + 3u;
            atomicAdd(&P.ctr[MF_CS * (C_ADM)], 1ull);
        }
        for (int64_t j = c.lane; j < L; j += MF_WAVE) {
            int32_t x1 = r.at(j);
            if (x1 < 0 || x1 == z) continue;
            if (MF_FA_T > 0 && row_len(g, x1) <= MF_FA_T) continue;
            // one returning exchange tells the first toucher of the batch;
            // the count is a no-return add (the allocation resets it)
            const bool first = atomicExch(&g.fa_tag[x1], P.b) != P.b;
            atomicAdd(&g.fa_n[x1], 1u);
            wave_push(P.big, &P.ctr[MF_CS * (C_BIG)], x1, first);
        }
    }
}

// P9: reserve index space per long row (no space: the row is scanned).  One
// reservation per wave on the pool top: returning atomics on one address
// serialise.
__device__ void ph_fa_alloc(const BP& P, const GC& c) {
    int64_t nb = (int64_t)P.ctr[MF_CS * (C_BIG)];
    bool ovf = nb > P.big_cap;
    if (ovf) nb = P.big_cap;
    // whole waves step together so the reservation below is wave-wide
    for (int64_t i0 = c.tid - c.lane; i0 < nb; i0 += c.nthreads) {
        const int64_t i = i0 + c.lane;
        int32_t x1 = -1;
        int64_t cnt = 0;
        if (i < nb) {
            x1 = P.big[i];
            cnt = (int64_t)P.g.fa_n[x1];
        }
        const int64_t incl = wave_incl_scan(cnt);
        const int64_t total = __shfl(incl, MF_WAVE - 1, MF_WAVE);
        u64 base = 0;
        if (c.lane == 0 && total) base = atomicAdd(&P.ctr[MF_CS * (C_FA_TOP)], (u64)total);
        base = (u64)__shfl((long long)base, 0, MF_WAVE);
        if (x1 < 0) continue;
        const int64_t s = (int64_t)base + (incl - cnt);
        P.g.fa_n[x1] = 0;  // zero again for the next batch's count
        P.g.fa_start[x1] = (ovf || s + cnt > P.fa_cap) ? -1 : s;
        P.g.fa_fill[x1] = 0;
    }
}

// P10: fill the index: (z, position of z in row(x1)).  ADM members are few
// and the grid has waves to spare, so a wave takes eight neighbours x1 of
// one ADM member z at a time and each x1 is searched by a group of 8 lanes
// (group_find: 6 dependent loads for a row of 2^17 instead of the 17 of one
// lane's binary search), then x1's overlay run when the base entry is absent
// or deleted (the flat position of row_pos).
#define MF_FA_PARTS 16  // waves per ADM member, eight neighbours each per step
__device__ void ph_fa_fill(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    constexpr int G = 8;
    const int grp = c.lane / G, gl = c.lane % G, gbase = grp * G;
    int64_t np = (int64_t)P.ctr[MF_CS * (C_PU)];
    if (np > P.pu_cap) np = P.pu_cap;
    for (int64_t t = c.wave; t < np * MF_FA_PARTS; t += c.nwaves) {
        const int32_t z = P.pu[t / MF_FA_PARTS];
        if (!is_adm(P, z)) continue;
        Row r = get_row(g, z);
        const int64_t L = r.len();
        for (int64_t j0 = (t % MF_FA_PARTS) * G; j0 < L; j0 += (int64_t)MF_FA_PARTS * G) {
            const int64_t j = j0 + grp;  // this group's neighbour: the 8 lanes agree
            int32_t x1 = -1;
            if (j < L) {
                const int32_t y = r.at(j);
                if (y >= 0 && y != z && (MF_FA_T == 0 || row_len(g, y) > MF_FA_T) && g.fa_start[y] >= 0) x1 = y;
            }
            if (x1 < 0) continue;
            const int64_t p = group_find<G>(g, x1, z, gl, gbase);
            int64_t pos = -1;
            if (p >= 0 && g.nlist[p] >= 0) {
                pos = p - g.nindex[x1];
            } else {
                const int32_t al = g.ov_len[x1];
                const int32_t* run = g.pool + g.ov_start[x1];
                for (int32_t t0 = 0; t0 < al; t0 += G) {
                    const int32_t q = t0 + gl;
                    const u64 hit = (__ballot(q < al && run[q] == z) >> gbase) & ((1ull << G) - 1ull);
                    if (hit) {
                        pos = (g.nindex[x1 + 1] - g.nindex[x1]) + t0 + __ffsll((long long)hit) - 1;
                        break;
                    }
                }
            }
            if (gl == 0 && pos >= 0) {
                const int64_t st = g.fa_start[x1];
                const int32_t slot = atomicAdd(&g.fa_fill[x1], 1);
                P.fa_z[st + slot] = z;
                P.fa_pos[st + slot] = (int32_t)pos;
            }
        }
    }
}

// ---------------------------------------------------------------- the rule
// Every outcome records only the reads it rests on.  A blocked outcome rests
// on one blocker (its witness): while that read holds, the outcome holds,
// whatever else changes.  A clear outcome (an admission) rests on every read
// of its scan, which is then walked a second time to record them.  Blocks
// by per-batch static facts (a vertex outside PU, no row, an MIS member
// outside W, which nothing can evict) need no record, and scans prefer such a
// blocker when they have one.

// Is x (an ADM vertex, x != w) admissible right after w's eviction?  xadj:
// x is a neighbour of w (w is one of the MIS neighbours x counts).  On false,
// `wit` is the blocking read; with ALL, every read is recorded.
//
// mx = mc[x] and sx = seedb[x], loaded by the caller together with x's PU
// tag (independent loads issued at once shorten the evaluation's chain).
template <bool ALL>
__device__ bool cand_ok_v(const BP& P, int32_t x, int32_t w, bool xadj, uint32_t mx, uint32_t sx, Rd& wit,
                          Recorder& rc) {
    const DevGraph& g = P.g;
    if (mx & MF_MISBIT) {
        wit = rd(x, mx, MF_T_MIS);
        return false;
    }
    uint32_t cx = mx & MF_CMASK;
    uint32_t corr = (xadj && cx) ? 1u : 0u;
    if (cx - corr) {
        wit = rd(x, mx, (1u + corr) | MF_T_NOFLIP);
        return false;
    }
    if (ALL) record(P, rc, rd(x, mx, 0u));
    if (sx != P.b) return true;  // candidate row = adj(x)
    // seed row {x} ∪ N1(x) ∪ N2(x) must hold no MIS member
    Row r = get_row(g, x);
    int64_t L = r.len();
    for (int64_t j = 0; j < L; j++) {
        int32_t y = r.at(j);
        if (y < 0 || y == x) continue;
        const uint32_t py = g.pub[y], my = g.mc[y], wy = g.wpotb[y];
        if (y == w) {  // w is outside the MIS now; its own count is unchanged
            if (my & MF_CMASK) {
                wit = rd(w, my, 1u | MF_T_NOFLIP);
                return false;
            }
            if (ALL) record(P, rc, rd(w, my, 0u));
            continue;
        }
        if ((py >> 2) != P.b) {  // outside PU
            wit = rd_none();
            return false;
        }
        if (my & MF_MISBIT) {
            wit = wy == P.b ? rd(y, my, MF_T_MIS) : rd_none();
            return false;
        }
        uint32_t cy = my & MF_CMASK;
        if (cy >= 2) {
            wit = rd(y, my, 2u | MF_T_NOFLIP);
            return false;
        }
        if (cy == 1 && !adjacent(g, y, w)) {  // its one MIS neighbour is not w
            wit = rd(y, my, 1u | MF_T_NOFLIP);
            return false;
        }
        if (ALL) record(P, rc, rd(y, my, 0u));
    }
    return true;
}

// The marks of a speculated flip's neighbours (mark_vertex), issued by an
// evaluation's pass over the flipped vertex's row from the counts it loads
// anyway: an evicted w's in part 2 of eval_insertion (`done`: the prefix of
// the row covered), an admitted deletion endpoint's in eval_deletion.
// A short row's marks are issued by the caller after the evaluation's last
// load (`dm`: this lane's neighbour `dx` is to be marked).
struct SpecMarks {
    u64 mv;
    uint32_t R;
    u64 bound;
    int64_t done;
    bool dm = false;
    int32_t dx = -1;
};

// Insertion with both ends in the MIS, w evicted: first admissible vertex of
// row(w) in BFS order, or -1.  Wave-cooperative; all lanes return the same.
// With `em`, w's neighbours are marked as part 2 passes them.  pw, mw and rw
// are w's PU word, count word and row, loaded by the caller (the
// speculation issues them with the edge's classification).
__device__ int32_t eval_insertion_from(const BP& P, int32_t w, uint32_t pw, uint32_t mw, const Row& rw, Recorder& rc,
                                       SpecMarks* em) {
    const DevGraph& g = P.g;
    const int lane = lane_id();
    Recorder none{REC_NONE};
    const int64_t L = rw.len();
    const bool fa_ok = P.ctr[MF_CS * (C_FA_OVF)] == 0;
    // position 0: w itself.  Outside PU, w keeps an MIS neighbour whenever it
    // is outside the MIS (its insertion partner), so it is blocked for sure.
    if ((pw >> 2) == P.b) {
        if (mw & MF_CMASK) {
            if (lane == 0) record(P, rc, rd(w, mw, 1u | MF_T_NOFLIP));
        } else {
            const Row& r = rw;
            bool blocked = false;
            for (int64_t j0 = 0; j0 < L && !blocked; j0 += MF_WAVE) {
                int64_t j = j0 + lane;
                bool bl = false, fixed = false;
                Rd wt = rd_none();
                if (j < L) {
                    int32_t y = r.at(j);
                    if (y >= 0 && y != w) {
                        const uint32_t py = g.pub[y], my = g.mc[y], wy = g.wpotb[y];
                        if ((py >> 2) != P.b) bl = fixed = true;  // outside PU
                        else {
                            if (my & MF_MISBIT) {
                                bl = true;
                                fixed = wy != P.b;  // a member outside W stays one
                                wt = rd(y, my, MF_T_MIS);
                            } else if ((my & MF_CMASK) >= 2) {  // w was one of them
                                bl = true;
                                wt = rd(y, my, 2u | MF_T_NOFLIP);
                            }
                        }
                    }
                }
                u64 bal = __ballot(bl);
                if (bal) {
                    blocked = true;
                    if (!__ballot(fixed) && lane == first_lane(bal)) record(P, rc, wt);
                }
            }
            if (!blocked) {  // w itself is admissible again: record every read
                if (lane == 0) record(P, rc, rd(w, mw, 0u));
                for (int64_t j = lane; j < L; j += MF_WAVE) {
                    int32_t y = r.at(j);
                    if (y >= 0 && y != w) record(P, rc, rd(y, g.mc[y], 0u));
                }
                return w;
            }
        }
    }
    // part 2: N1(w) in adjacency order.  A row of at most one wave's length
    // (the usual case) also loads each neighbour's index words here, with its
    // other words, for part 3's pass over the same neighbours.
    const bool short_row = L <= MF_WAVE;
    int32_t sx1 = -1, sfill = 0;
    uint32_t stag = 0;
    int64_t sstart = -1;
    for (int64_t j0 = 0; j0 < L; j0 += MF_WAVE) {
        int64_t j = j0 + lane;
        int32_t x = -1;
        bool cnd = false, ok = false;
        Rd wt = rd_none();
        uint32_t mx = 0, sx = 0;
        if (j < L) {
            x = rw.at(j);
            if (x >= 0 && x != w) {
                const uint32_t px = g.pub[x];
                mx = g.mc[x];
                sx = g.seedb[x];
                if (short_row) {
                    sx1 = x;
                    stag = g.fa_tag[x];
                    sstart = g.fa_start[x];
                    sfill = g.fa_fill[x];
                }
                if (em && (u64)(mx & MF_CMASK) < 2ull + em->bound) {
                    if (short_row) {
                        em->dm = true;
                        em->dx = x;
                    } else {
                        atomicMax(&g.cmark[x], em->mv);
                        round_count_add(&g.cdec[x], em->R);
                    }
                }
                if (px == 4u * P.b + 3u) {  // admissible
                    cnd = true;
                    ok = cand_ok_v<false>(P, x, w, true, mx, sx, wt, none);
                }
            }
        }
        if (em) em->done = j0 + MF_WAVE < L ? j0 + MF_WAVE : L;
        u64 bal = __ballot(ok);
        const int f = bal ? first_lane(bal) : MF_WAVE;
        if (cnd && !ok && lane < f) record(P, rc, wt);  // earlier candidates stay blocked
        if (bal) {
            if (lane == f) (void)cand_ok_v<true>(P, x, w, true, mx, sx, wt, rc);  // the words it read above
            return __shfl(x, f, MF_WAVE);
        }
    }
    // part 3: concat over x1 in adj(w) of adj(x1); long rows through the index
    for (int64_t j0 = 0; j0 < L; j0 += MF_WAVE) {
        int64_t j = j0 + lane;
        int kindl = 0;  // 0 empty, 1 row scan, 2 index
        const int32_t* pb = nullptr;
        const int32_t* pa = nullptr;
        int64_t blen = 0, slen = 0;
        if (j < L) {
            int32_t x1 = short_row ? sx1 : rw.at(j);
            if (x1 >= 0 && x1 != w) {
                // every row with an admissible neighbour has index entries
                // this batch; only an overflowed index falls back to the row
                bool indexed = (short_row ? stag : g.fa_tag[x1]) == P.b;
                int64_t st = indexed && fa_ok ? (short_row ? sstart : g.fa_start[x1]) : -1;
                if (st >= 0) {
                    kindl = 2;
                    pb = P.fa_z + st;
                    pa = P.fa_pos + st;
                    slen = short_row ? sfill : g.fa_fill[x1];
                } else if (indexed || !fa_ok) {
                    Row r1 = get_row(g, x1);
                    kindl = 1; pb = r1.base; pa = r1.add; blen = r1.blen; slen = r1.len();
                }
            }
        }
        int64_t incl = wave_incl_scan(slen);
        int64_t excl = incl - slen;
        int64_t total = __shfl(incl, MF_WAVE - 1, MF_WAVE);
        u64 best = ~0ull;
        int32_t best_x = -1;
        uint32_t best_m = 0, best_s = 0;  // the best candidate's count and seed words
        for (int64_t k0 = 0; k0 < total; k0 += MF_WAVE) {
            int64_t k = k0 + lane;
            int64_t kk = k < total ? k : total - 1;
            // owner lane of flat index kk: largest lane whose segment starts
            // at or before kk (fixed 6 halvings, all lanes active)
            int lo = 0, hi = MF_WAVE - 1;
#pragma unroll
            for (int step = 0; step < 6; step++) {
                int mid = (lo + hi + 1) >> 1;
                int64_t om = __shfl(excl, mid, MF_WAVE);
                if (om <= kk) lo = mid; else hi = mid - 1;
            }
            const int owner = lo;
            int64_t oexcl = __shfl(excl, owner, MF_WAVE);
            int okind = __shfl(kindl, owner, MF_WAVE);
            const int32_t* obase = (const int32_t*)(intptr_t)__shfl((long long)(intptr_t)pb, owner, MF_WAVE);
            const int32_t* oadd = (const int32_t*)(intptr_t)__shfl((long long)(intptr_t)pa, owner, MF_WAVE);
            int64_t oblen = __shfl(blen, owner, MF_WAVE);
            u64 key = ~0ull;
            int32_t x = -1;
            uint32_t px = 0, mx = 0, sx = 0;
            if (k < total) {
                int64_t off = k - oexcl;
                int64_t pos;
                if (okind == 2) {
                    x = obase[off];
                    pos = oadd[off];
                } else {
                    x = off < oblen ? obase[off] : oadd[off - oblen];
                    pos = off;
                }
                if (x >= 0 && x != w) {
                    px = g.pub[x];
                    mx = g.mc[x];
                    sx = g.seedb[x];
                }
                if (x >= 0 && x != w && px == 4u * P.b + 3u) {  // admissible
                    Rd wt = rd_none();
                    if (cand_ok_v<false>(P, x, w, false, mx, sx, wt, none))
                        key = ((u64)(uint32_t)owner << 32) | (u64)(uint32_t)pos;
                    else
                        record(P, rc, wt);  // (also for candidates after the result: harmless)
                }
            }
            u64 mn = wave_min_u64(key);
            if (mn < best) {
                u64 bal = __ballot(key == mn);
                const int fl = first_lane(bal);
                best = mn;
                best_x = __shfl(x, fl, MF_WAVE);
                best_m = (uint32_t)__shfl((int)mx, fl, MF_WAVE);
                best_s = (uint32_t)__shfl((int)sx, fl, MF_WAVE);
            }
            if (best != ~0ull) {
                // stop once every element of the best owner's segment is seen
                int bo = (int)(best >> 32);
                int64_t end_bo = __shfl(incl, bo, MF_WAVE);
                if (k0 + MF_WAVE >= end_bo) break;
            }
        }
        if (best != ~0ull) {
            if (lane == 0) {
                Rd wt;
                (void)cand_ok_v<true>(P, best_x, w, false, best_m, best_s, wt, rc);
            }
            return best_x;
        }
    }
    return -1;
}

__device__ __forceinline__ int32_t eval_insertion(const BP& P, int32_t w, Recorder& rc, SpecMarks* em = nullptr) {
    return eval_insertion_from(P, w, P.g.pub[w], P.g.mc[w], get_row(P.g, w), rc, em);
}

// Deletion, exactly one end `other` in the MIS: may tgt join?  row(tgt) =
// {tgt} ∪ N1 ∪ N2 must hold no MIS member other than `other`.
// With `em` (its `done` unused), the admission marks of tgt's neighbours
// (mark_vertex) are issued by the final pass that records the reads.
__device__ bool eval_deletion_from(const BP& P, int32_t tgt, int32_t other, const Row& r, Recorder& rc,
                                   const SpecMarks* em) {
    const DevGraph& g = P.g;
    const int lane = lane_id();
    int64_t L = r.len();
    for (int64_t j0 = 0; j0 < L; j0 += MF_WAVE) {
        int64_t j = j0 + lane;
        bool bl = false, fixed = false;
        Rd wt = rd_none();
        if (j < L) {
            int32_t y = r.at(j);
            if (y >= 0 && y != tgt) {
                const uint32_t my = g.mc[y], wy = g.wpotb[y];
                uint32_t cy = my & MF_CMASK;
                if (y == other) {  // its MIS neighbours sit two hops from tgt
                    if (cy) { bl = true; wt = rd(y, my, 1u | MF_T_NOFLIP); }
                } else if (my & MF_MISBIT) {
                    bl = true;
                    fixed = wy != P.b;  // a member outside W stays one
                    wt = rd(y, my, MF_T_MIS);
                } else if (cy >= 2) {
                    bl = true;
                    wt = rd(y, my, 2u | MF_T_NOFLIP);
                } else if (cy == 1 && !adjacent(g, y, other)) {
                    bl = true;
                    wt = rd(y, my, 1u | MF_T_NOFLIP);
                }
            }
        }
        u64 bal = __ballot(bl);
        if (bal) {
            if (!__ballot(fixed) && lane == first_lane(bal)) record(P, rc, wt);
            return false;
        }
    }
    // tgt joins: every read of the scan decides it
    for (int64_t j = lane; j < L; j += MF_WAVE) {
        int32_t y = r.at(j);
        if (y >= 0 && y != tgt) {
            const uint32_t my = g.mc[y];
            record(P, rc, rd(y, my, 0u));
            if (em && (u64)(my & MF_CMASK) < 2ull + em->bound) atomicMax(&g.cmark[y], em->mv);
        }
    }
    return true;
}

__device__ __forceinline__ bool eval_deletion(const BP& P, int32_t tgt, int32_t other, Recorder& rc,
                                              const SpecMarks* em = nullptr) {
    return eval_deletion_from(P, tgt, other, get_row(P.g, tgt), rc, em);
}

// ---------------------------------------------------------------- rounds
// The round state lives on the device: ctr[C_START0 + par] is the first edge
// not yet committed, read by the rounds of parity par, and *P.round the id of
// the last round.  A round whose start is at or past U does nothing, so
// rounds can be queued without knowing how many the batch needs; only the
// commit phase advances the state, and it writes the next round's start into
// the other parity's word, so no wave of the committing kernel can see a
// half-advanced round.
struct RoundView {
    int64_t start, end;
    uint32_t R;
    bool live;
};

// `win` bounds the edges a round speculates: one per wave of the round grid
// for the grid-wide rounds (P.window), MF_FIN_WINDOW in the one-workgroup
// tail, so a round costs about the same whatever the batch size.
#define MF_FIN_WINDOW 1024
// A tail that may halt leaves more than this many uncommitted edges to
// further grid-wide rounds queued by the host: one workgroup is far slower
// per round than the whole GPU.
#define MF_HALT_EDGES 256
__device__ __forceinline__ RoundView round_view(const BP& P, int par, int64_t win) {
    RoundView v;
    v.start = (int64_t)P.ctr[MF_CS * (C_START0 + par)];
    v.end = v.start + win < P.U ? v.start + win : P.U;
    v.R = *P.round + 1u;
    v.live = v.start < P.U;
    return v;
}

// Kind of an edge from its endpoints' count words: 1 an insertion with both
// ends in the MIS, 2 a deletion with exactly one end in it, 0 inactive.
__device__ __forceinline__ uint8_t edge_kind(bool ins, uint32_t ma, uint32_t mb) {
    const bool ia = (ma & MF_MISBIT) != 0, ib = (mb & MF_MISBIT) != 0;
    if (ins) return (ia && ib) ? 1 : 0;
    return (ia != ib) ? 2 : 0;
}

__device__ __forceinline__ uint8_t classify_edge(const BP& P, bool ins, int32_t a, int32_t bb) {
    return edge_kind(ins, P.g.mc[a], P.g.mc[bb]);
}

// What the evaluation of an edge needs first that depends on its endpoints
// alone, loaded together with the classification's count words: an
// insertion evaluates row(max(a, b)), a deletion the row of the end outside
// the MIS, so both rows' bounds come with the PU word of the larger end.
struct EdgeLoads {
    int32_t a, b;
    uint32_t ma, mb, pw;
    Row ra, rb;
};
__device__ __forceinline__ EdgeLoads edge_loads(const BP& P, int32_t a, int32_t b) {
    EdgeLoads l;
    l.a = a;
    l.b = b;
    l.ma = P.g.mc[a];
    l.mb = P.g.mc[b];
    l.pw = P.g.pub[a > b ? a : b];
    l.ra = get_row(P.g, a);
    l.rb = get_row(P.g, b);
    return l;
}

// Marks the writes of a speculated membership flip of z: z's bit and the
// counts of its neighbours (evictions also raise the neighbours' decrement
// bound for this round).  A neighbour whose count is at least 2 + `bound`
// (bound >= evictions this round) is left unmarked: no read of it can
// change its outcome this round (every count threshold is at most 2).
// Entries before `from` were marked already (by the evaluation's first pass
// over an evicted vertex's row, SpecMarks).
// `pre`: z's row, when the caller has it.
__device__ __forceinline__ void mark_vertex(const BP& P, int32_t z, u64 mv, bool evict, uint32_t R, u64 bound,
                                            int lane, int64_t from = 0, const Row* pre = nullptr) {
    if (lane == 0) atomicMax(&P.g.mmark[z], mv);
    const Row r = pre ? *pre : get_row(P.g, z);
    int64_t L = r.len();
    for (int64_t j = from + lane; j < L; j += MF_WAVE) {
        int32_t y = r.at(j);
        if (y >= 0 && y != z) {
            if ((u64)(P.g.mc[y] & MF_CMASK) >= 2ull + bound) continue;
            atomicMax(&P.g.cmark[y], mv);
            if (evict) round_count_add(&P.g.cdec[y], R);
        }
    }
}

// Speculative evaluation of active edge e (kind k) by one wave, from its
// endpoints' loads; the reads its outcome rests on go to `rc` (nowhere, or
// the wave's read log).  Loads come first, then the marks, the results and
// the log records (stores and atomics ahead of a load delay its use).
__device__ __forceinline__ void spec_edge_from(const BP& P, int32_t e, uint8_t k, uint32_t R, u64 bound, int lane,
                                               Recorder& rc, int4* cmw, const EdgeLoads& el) {
    const int32_t a = el.a, bb = el.b;
    rc.R = R;
    rc.e = e;
    u64 mv = ((u64)R << 32) | (u64)(0xFFFFFFFFu - (uint32_t)e);
    if (k == 1) {
        const bool wa = a > bb;
        const int32_t w = wa ? a : bb;
        const Row rw = wa ? el.ra : el.rb;
        SpecMarks em{mv, R, bound, 0};
        int32_t x = eval_insertion_from(P, w, el.pw, wa ? el.ma : el.mb, rw, rc, &em);
        if (x >= 0) mark_vertex(P, x, mv, false, R, bound, lane);
        mark_vertex(P, w, mv, true, R, bound, lane, em.done, &rw);
        if (em.dm) {
            atomicMax(&P.g.cmark[em.dx], mv);
            round_count_add(&P.g.cdec[em.dx], R);
        }
        if (lane == 0) {
            P.res_w[e] = w;
            P.res_x[e] = x;
            if (cmw) *cmw = make_int4(1, w, x, e);
        }
    } else {
        const bool ma = (el.ma & MF_MISBIT) != 0;
        int32_t tgt = ma ? bb : a, other = ma ? a : bb;
        const SpecMarks em{mv, R, bound, 0};
        bool ok = eval_deletion_from(P, tgt, other, ma ? el.rb : el.ra, rc, &em);
        if (lane == 0) {
            if (ok) atomicMax(&P.g.mmark[tgt], mv);  // its neighbours were marked by the evaluation
            P.res_w[e] = -1;
            P.res_x[e] = ok ? tgt : -1;
            if (cmw) *cmw = make_int4(2, -1, ok ? tgt : -1, e);
        }
    }
    flush_records(rc);
}

__device__ __forceinline__ void spec_edge(const BP& P, int32_t e, uint8_t k, uint32_t R, u64 bound, int lane,
                                          Recorder& rc, int4* cmw) {
    spec_edge_from(P, e, k, R, bound, lane, rc, cmw, edge_loads(P, P.bu[e], P.bv[e]));
}

// Re-walk of active edge e: true when a read its outcome rests on may have
// been changed by an earlier edge of the round.  Wave-uniform.
__device__ __forceinline__ bool validate_edge(const BP& P, int32_t e, uint8_t k, uint32_t R) {
    int32_t a = P.bu[e], bb = P.bv[e];
    Recorder rc{REC_CHECK};
    rc.R = R;
    rc.e = e;
    rc.conf = mk_before(P.g.mmark, a, R, e) || mk_before(P.g.mmark, bb, R, e);
    if (k == 1) {
        int32_t w = a > bb ? a : bb;
        (void)eval_insertion(P, w, rc);
    } else {
        bool ma = (P.g.mc[a] & MF_MISBIT) != 0;
        int32_t tgt = ma ? bb : a, other = ma ? a : bb;
        (void)eval_deletion(P, tgt, other, rc);
    }
    return __ballot(rc.conf) != 0;
}

__device__ __forceinline__ void apply_flip(const BP& P, int32_t z, bool join, int lane) {
    if (lane == 0) {
        if (join) atomicOr(&P.g.mc[z], MF_MISBIT);
        else atomicAnd(&P.g.mc[z], MF_CMASK);
    }
    Row r = get_row(P.g, z);
    int64_t L = r.len();
    for (int64_t j = lane; j < L; j += MF_WAVE) {
        int32_t y = r.at(j);
        if (y >= 0 && y != z) {
            if (join) atomicAdd(&P.g.mc[y], 1u);
            else {
                atomicSub(&P.g.mc[y], 1u);
                P.g.nevb[y] = P.b;
            }
        }
    }
}

// An edge's flips: the evicted w leaves (k == 1), the admitted x joins (x >=
// 0).  Both rows' bounds and first wave of entries are loaded before any
// atomic is issued; longer rows continue as apply_flip does.
__device__ __forceinline__ void commit_flips(const BP& P, uint8_t k, int32_t w, int32_t x, int lane) {
    const DevGraph& g = P.g;
    const bool ev = k == 1, ad = x >= 0;
    Row rw{}, rx{};
    if (ev) rw = get_row(g, w);
    if (ad) rx = get_row(g, x);
    const int64_t Lw = ev ? rw.len() : 0, Lx = ad ? rx.len() : 0;
    const int32_t yw = lane < Lw ? rw.at(lane) : -1;
    const int32_t yx = lane < Lx ? rx.at(lane) : -1;
    if (lane == 0) {
        if (ev) {
            atomicAnd(&g.mc[w], MF_CMASK);
            g.evb[w] = P.b;
        }
        if (ad) atomicOr(&g.mc[x], MF_MISBIT);
    }
    if (yw >= 0 && yw != w) {
        atomicSub(&g.mc[yw], 1u);
        g.nevb[yw] = P.b;
    }
    if (yx >= 0 && yx != x) atomicAdd(&g.mc[yx], 1u);
    for (int64_t j = MF_WAVE + lane; j < Lw; j += MF_WAVE) {
        const int32_t y = rw.at(j);
        if (y >= 0 && y != w) {
            atomicSub(&g.mc[y], 1u);
            g.nevb[y] = P.b;
        }
    }
    for (int64_t j = MF_WAVE + lane; j < Lx; j += MF_WAVE) {
        const int32_t y = rx.at(j);
        if (y >= 0 && y != x) atomicAdd(&g.mc[y], 1u);
    }
}

__device__ __forceinline__ void commit_edge(const BP& P, int32_t e, uint8_t k, int lane) {
    commit_flips(P, k, k == 1 ? P.res_w[e] : -1, P.res_x[e], lane);
}

// First conflicting edge of the round (INT64_MAX when none): the counter
// holds ~0 until some edge conflicts.
__device__ __forceinline__ int64_t round_conflict(const BP& P, int par) {
    u64 cf = P.ctr[MF_CS * (C_CONF0 + par)];
    return cf > (u64)INT64_MAX ? INT64_MAX : (int64_t)cf;
}

// Round bookkeeping after a commit (one thread): the next round (parity
// par ^ 1) starts at the first conflict.  A dead round carries the start
// over unchanged.
__device__ __forceinline__ void advance_round(const BP& P, int par, int64_t conflict, int64_t win) {
    RoundView v = round_view(P, par, win);
    P.ctr[MF_CS * (C_NACT0 + (par ^ 1))] = 0;
    P.ctr[MF_CS * (C_CONF0 + (par ^ 1))] = ~0ull;
    if (v.live) {
        P.ctr[MF_CS * (C_ROUNDS)] += 1;
        P.ctr[MF_CS * (C_SPECS)] += (u64)(v.end - v.start);
        P.ctr[MF_CS * (C_START0 + (par ^ 1))] = (u64)(conflict < v.end ? conflict : v.end);
        *P.round = v.R;
    } else {
        P.ctr[MF_CS * (C_START0 + (par ^ 1))] = (u64)v.start;
    }
}

// Grid-wide rounds: one wave per edge of the window.  Speculation classifies
// the edge first (the kind depends only on the committed state) and logs its
// guarded reads, so a round is three kernels: speculate, validate, commit.
// `cur` is the LDS cursor array of the workgroup's waves.  `st` is the
// wave's staged edge (stage[wave], loaded with the round's start): when it
// is the wave's first edge, its endpoints need no load after the start's.
// The segment's header (slot 0) gets the record count and the kind (bits 16
// and up), the index and the endpoints of the wave's first edge (index
// 0xFFFFFFFF: none), so the validation needs no load of its own for them.
#define MF_NO_EDGE 0xFFFFFFFFu
__device__ void ph_spec_w(const BP& P, const GC& c, const RoundView& v, int* cur, int4 st) {
    const int wiw = (int)(threadIdx.x / MF_WAVE);
    if (c.lane == 0) cur[wiw] = 0;
    __syncthreads();
    Recorder rc{REC_LOG, ReadLog{P.rlog + c.wave * MF_RLOG, cur + wiw, P.relog}};
    uint4 hdr = make_uint4(0u, MF_NO_EDGE, 0u, 0u);
    for (int64_t e = v.start + c.wave; e < v.end; e += c.nwaves) {
        const bool first = e == v.start + c.wave;
        int32_t a, bb;
        bool ins;
        if (first && st.w == (int32_t)e && ((uint32_t)st.z & ~1u) == (P.b << 1)) {  // this batch's edge e
            a = st.x;
            bb = st.y;
            ins = (st.z & 1) != 0;
        } else {
            a = P.bu[e];
            bb = P.bv[e];
            ins = P.bins[e] != 0;
        }
        const EdgeLoads el = edge_loads(P, a, bb);
        const uint8_t k = edge_kind(ins, el.ma, el.mb);
        // the commit's record of the wave's first edge (wave w of the commit
        // commits edge start + w)
        int4* cmw = first ? P.cm + c.wave : nullptr;
        if (first) hdr = make_uint4((uint32_t)k << 16, (uint32_t)e, (uint32_t)a, (uint32_t)bb);
        if (k) spec_edge_from(P, (int32_t)e, k, v.R, (u64)P.n_ins, c.lane, rc, cmw, el);
        if (c.lane == 0) {  // after the evaluation's loads
            P.kind[e] = k;
            if (cmw && !k) *cmw = make_int4(0, -1, -1, (int32_t)e);
        }
    }
    if (c.lane == 0) {
        const int n = cur[wiw];
        hdr.x |= (uint32_t)(n < MF_RLOG - 1 ? n : MF_RLOG - 1);
        P.rlog[c.wave * MF_RLOG] = hdr;
    }
}

// Validation: wave w checks what wave w of the speculation (same grid) did:
// the reads logged in its segment rlog[w * MF_RLOG ...] and its edge's
// endpoints (the classification read them), with a re-walk when the log
// overflowed.  A wave with no edge in the window reads nothing.  One load
// of the segment's first line gives the count and up to seven records
// (most segments hold only a few), issued together with the edge's own
// loads; the rest of a longer segment is loaded when the count asks for it.
static_assert(MF_RLOG == 2 * MF_WAVE, "two log slots per lane");
__device__ __forceinline__ bool endpoint_conflict(const BP& P, int64_t e, uint8_t k, int32_t a, int32_t bb,
                                                  uint32_t R) {
    if (k && P.relog[e] == R) return validate_edge(P, (int32_t)e, k, R);
    return mk_before(P.g.mmark, a, R, (int32_t)e) || mk_before(P.g.mmark, bb, R, (int32_t)e);
}

#define MF_RLOG_LINE 8  // 16-byte slots per 128-byte line
// `r0`: slot `lane` of the wave's segment (lanes below MF_RLOG_LINE), loaded
// with the round's start.
__device__ void ph_validate_w(const BP& P, const GC& c, const RoundView& v, int par, uint4 r0) {
    const uint4* seg = P.rlog + c.wave * MF_RLOG;
    const int64_t e0 = v.start + c.wave;
    if (e0 >= v.end) return;  // wave-uniform: no edge, no records
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t hx = (uint32_t)__shfl((int)r0.x, 0, MF_WAVE);
    const uint32_t he = (uint32_t)__shfl((int)r0.y, 0, MF_WAVE);
    uint8_t k0;
    int32_t a0, b0;
    if (he == (uint32_t)e0) {  // the header names the wave's edge (always, with the speculation's grid)
        k0 = (uint8_t)(hx >> 16);
        a0 = __shfl((int)r0.z, 0, MF_WAVE);
        b0 = __shfl((int)r0.w, 0, MF_WAVE);
    } else {
        k0 = P.kind[e0];
        a0 = P.bu[e0];
        b0 = P.bv[e0];
    }
    const int n = (int)(hx & 0xFFFFu);
    // slot l (1 <= l <= n) is checked by lane l, slot 64 + l by lane l
    if (c.lane < MF_RLOG_LINE) {
        if (c.lane > n) r0 = zero;
    } else {
        r0 = c.lane <= n ? seg[c.lane] : zero;
    }
    uint4 r1 = zero;
    if (c.lane + MF_WAVE <= n) r1 = seg[c.lane + MF_WAVE];
    bool c0 = false, c1 = false;
    if (c.lane >= 1 && c.lane <= n) vread(P, (int32_t)r0.x, r0.y, r0.z, v.R, (int32_t)r0.w, c0);
    if (c.lane + MF_WAVE <= n) vread(P, (int32_t)r1.x, r1.y, r1.z, v.R, (int32_t)r1.w, c1);
    u64 mine = ~0ull;
    if (c0) mine = (u64)r0.w;
    if (c1 && (u64)r1.w < mine) mine = (u64)r1.w;
    if (endpoint_conflict(P, e0, k0, a0, b0, v.R) && (u64)e0 < mine) mine = (u64)e0;  // wave-uniform call
    // a window wider than the grid gives a wave further edges
    for (int64_t e = e0 + c.nwaves; e < v.end; e += c.nwaves)
        if (endpoint_conflict(P, e, P.kind[e], P.bu[e], P.bv[e], v.R) && (u64)e < mine) mine = (u64)e;
    mine = wave_min_u64(mine);
    if (mine != ~0ull && c.lane == 0) atomicMin(&P.ctr[MF_CS * (C_CONF0 + par)], mine);
}

// Wave w commits edge start + w from the record wave w of the speculation
// left (loaded beside the conflict index, not after it), further edges of a
// wider window from the arrays.  The grid also stages the next round's
// edges: stage[i] = edge next + i, where next (the first conflict, or the
// window's end) is where the next round starts, tagged with the batch epoch
// (z = epoch << 1 | insertion flag) and the edge's index.
__device__ void ph_commit_w(const BP& P, const GC& c, const RoundView& v, int par) {
    const int64_t e0 = v.start + c.wave;
    const int4 m = e0 < v.end ? P.cm[c.wave] : make_int4(0, -1, -1, -1);
    const int64_t conflict = round_conflict(P, par);
    const int64_t lim = conflict < v.end ? conflict : v.end;
    if (e0 < lim) {
        if (m.w == (int32_t)e0) {
            if (m.x) commit_flips(P, (uint8_t)m.x, m.y, m.z, c.lane);
        } else {  // no record for this edge (not expected: both kernels use the round grid)
            const uint8_t k = P.kind[e0];
            if (k) commit_edge(P, (int32_t)e0, k, c.lane);
        }
    }
    for (int64_t e = e0 + c.nwaves; e < lim; e += c.nwaves) {
        uint8_t k = P.kind[e];
        if (k) commit_edge(P, (int32_t)e, k, c.lane);
    }
    // after the commit's own loads
    const int64_t ns = P.window < P.nstage ? P.window : P.nstage;
    for (int64_t i = c.tid; i < ns; i += c.nthreads) {
        const int64_t e = lim + i;
        P.stage[i] = e < P.U ? make_int4(P.bu[e], P.bv[e], (int)((P.b << 1) | (uint32_t)(P.bins[e] != 0)), (int32_t)e)
                             : make_int4(-1, -1, 0, -1);
    }
    if (c.tid == 0) advance_round(P, par, conflict, P.window);
}

// One-workgroup rounds (finisher): classification by thread into an active
// list, then a wave per active edge, which suits the short tails it gets.
__device__ void ph_classify_l(const BP& P, const GC& c, const RoundView& v, int par) {
    for (int64_t e = v.start + c.tid; e < v.end; e += c.nthreads) {
        int32_t a = P.bu[e], bb = P.bv[e];
        uint8_t k = classify_edge(P, P.bins[e] != 0, a, bb);
        P.kind[e] = k;
        wave_push(P.active, &P.ctr[MF_CS * (C_NACT0 + par)], (int32_t)e, k != 0);
    }
}

__device__ void ph_spec_l(const BP& P, const GC& c, int par, uint32_t R) {
    int64_t na = (int64_t)P.ctr[MF_CS * (C_NACT0 + par)];
    for (int64_t i = c.wave; i < na; i += c.nwaves) {
        int32_t e = P.active[i];
        Recorder none{REC_NONE};
        spec_edge(P, e, P.kind[e], R, (u64)na, c.lane, none, nullptr);
    }
}

__device__ void ph_validate_l(const BP& P, const GC& c, const RoundView& v, int par) {
    const uint32_t R = v.R;
    for (int64_t e = v.start + c.tid; e < v.end; e += c.nthreads) {
        if (P.kind[e]) continue;
        int32_t a = P.bu[e], bb = P.bv[e];
        if (mk_before(P.g.mmark, a, R, (int32_t)e) || mk_before(P.g.mmark, bb, R, (int32_t)e))
            atomicMin(&P.ctr[MF_CS * (C_CONF0 + par)], (u64)e);
    }
    int64_t na = (int64_t)P.ctr[MF_CS * (C_NACT0 + par)];
    for (int64_t i = c.wave; i < na; i += c.nwaves) {
        int32_t e = P.active[i];
        if (validate_edge(P, e, P.kind[e], R) && c.lane == 0) atomicMin(&P.ctr[MF_CS * (C_CONF0 + par)], (u64)e);
    }
}

__device__ void ph_commit_l(const BP& P, const GC& c, int par) {
    int64_t na = (int64_t)P.ctr[MF_CS * (C_NACT0 + par)];
    const int64_t conflict = round_conflict(P, par);
    for (int64_t i = c.wave; i < na; i += c.nwaves) {
        int32_t e = P.active[i];
        if (e < conflict) commit_edge(P, e, P.kind[e], c.lane);
    }
}

// ---------------------------------------------------------------- completion
__device__ __forceinline__ void ph_comp_collect(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    int64_t np = (int64_t)P.ctr[MF_CS * (C_PU)];
    if (np > P.pu_cap) np = P.pu_cap;
    for (int64_t i = c.tid; i < np; i += c.nthreads) {
        int32_t z = P.pu[i];
        // outside the MIS with no MIS neighbour, and in C
        bool open = g.mc[z] == 0 && (g.seedb[z] == P.b || g.evb[z] == P.b || g.nevb[z] == P.b);
        if (open) g.cstat[z] = (P.b << 2) | 1u;
        wave_push(P.cand, &P.ctr[MF_CS * (C_CAND)], z, open);
    }
}

// A candidate joins when no open candidate neighbour has a smaller id
// (sequential greedy in ascending id, decided in parallel rounds).  One wave
// per candidate, so a long row is scanned 64 entries at a time.
__device__ void ph_comp_decide(const BP& P, const GC& c) {
    const DevGraph& g = P.g;
    int64_t nc = (int64_t)P.ctr[MF_CS * (C_CAND)];
    const uint32_t open = (P.b << 2) | 1u;
    for (int64_t i = c.wave; i < nc; i += c.nwaves) {
        int32_t z = P.cand[i];
        uint8_t d = 0;
        if (g.cstat[z] == open) {
            if (g.mc[z] & MF_CMASK) d = 3;
            else {
                d = 2;
                Row r = get_row(g, z);
                int64_t L = r.len();
                for (int64_t j0 = 0; j0 < L; j0 += MF_WAVE) {
                    int64_t j = j0 + c.lane;
                    bool sm = false;
                    if (j < L) {
                        int32_t y = r.at(j);
                        sm = y >= 0 && y < z && g.cstat[y] == open;
                    }
                    if (__ballot(sm)) { d = 0; break; }
                }
            }
        }
        if (c.lane == 0) P.cdec[i] = d;
    }
}

__device__ void ph_comp_apply(const BP& P, const GC& c, int par) {
    const DevGraph& g = P.g;
    int64_t nc = (int64_t)P.ctr[MF_CS * (C_CAND)];
    const uint32_t open = (P.b << 2) | 1u;
    for (int64_t i = c.wave; i < nc; i += c.nwaves) {
        int32_t z = P.cand[i];
        uint8_t d = P.cdec[i];
        if (d == 2) {
            apply_flip(P, z, true, c.lane);
            if (c.lane == 0) {
                g.cstat[z] = (P.b << 2) | 2u;
                atomicAdd(&P.ctr[MF_CS * (C_COMP_ADDS)], 1ull);
            }
        } else if (d == 3) {
            if (c.lane == 0) g.cstat[z] = (P.b << 2) | 3u;
        } else if (c.lane == 0 && g.cstat[z] == open) {
            atomicAdd(&P.ctr[MF_CS * (C_UNDEC0 + par)], 1ull);
        }
    }
    if (c.tid == 0) {
        P.ctr[MF_CS * (C_UNDEC0 + (par ^ 1))] = 0;
        P.ctr[MF_CS * (C_COMP_ROUNDS)] += 1;
    }
}

// ---------------------------------------------------------------- stats
// Decisions of committed edges stay in kind/res (later rounds only classify
// edges past the committed prefix), so the rule's counters are taken once
// per batch instead of with per-edge atomics.
#define MF_STATS_UNROLL 4  // quads per thread whose loads are all issued before the first is used
__device__ void ph_stats(const BP& P, const GC& c) {
    u64 ai = 0, ti = 0, ad = 0, td = 0;
    // four kinds per 32-bit load (kind is a whole allocation); every load of
    // a step is issued at a clamped index before any is used, so a thread
    // waits once per step instead of once per quad and again per result
    const int64_t nq = P.U / 4;
    const uint32_t* k4 = reinterpret_cast<const uint32_t*>(P.kind);
    for (int64_t q0 = c.tid; q0 < nq; q0 += MF_STATS_UNROLL * c.nthreads) {
        uint32_t w[MF_STATS_UNROLL];
        int32_t r[MF_STATS_UNROLL][4];
#pragma unroll
        for (int u = 0; u < MF_STATS_UNROLL; u++) {
            const int64_t q = q0 + u * c.nthreads;
            const int64_t qc = q < nq ? q : nq - 1;
            w[u] = k4[qc];
#pragma unroll
            for (int j = 0; j < 4; j++) r[u][j] = P.res_x[4 * qc + j];
            if (q >= nq) w[u] = 0;
        }
#pragma unroll
        for (int u = 0; u < MF_STATS_UNROLL; u++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t k = (w[u] >> (8 * j)) & 0xFFu;
                if (k == 1) { ai++; ti += r[u][j] >= 0; }
                else if (k == 2) { ad++; td += r[u][j] >= 0; }
            }
        }
    }
    for (int64_t e = 4 * nq + c.tid; e < P.U; e += c.nthreads) {
        uint8_t k = P.kind[e];
        if (k == 1) { ai++; ti += P.res_x[e] >= 0; }
        else if (k == 2) { ad++; td += P.res_x[e] >= 0; }
    }
    ai = wave_sum_u64(ai);
    if (ai && c.lane == 0) { atomicAdd(&P.ctr[MF_CS * (C_ACT_INS)], ai); atomicAdd(&P.ctr[MF_CS * (C_EVICT)], ai); }
    wave_add(&P.ctr[MF_CS * (C_ADMIT_INS)], ti);
    wave_add(&P.ctr[MF_CS * (C_ACT_DEL)], ad);
    wave_add(&P.ctr[MF_CS * (C_ADMIT_DEL)], td);
}

// ---------------------------------------------------------------- kernels
// A batch handed over in pinned host memory (P.hu set): its updates are
// read here into device arrays (9 bytes per update over PCIe, read rather
// than copied on another stream: a cross-stream wait costs more than these
// reads), and the batch start is stamped on the device clock.  A batch
// staged in device memory needs no such kernel: the claim kernel stamps the
// start.  The counters need no reset here: every batch's last kernel leaves
// them reset for the next (publish).
__device__ void ph_init(const BP& P, const GC& c) {
    if (c.tid == 0) P.ctr[MF_CS * (C_T0)] = wall_clock64();
    int32_t* du = const_cast<int32_t*>(P.bu);
    int32_t* dv = const_cast<int32_t*>(P.bv);
    uint8_t* di = const_cast<uint8_t*>(P.bins);
    for (int64_t i = c.tid; i < P.U; i += c.nthreads) {
        du[i] = P.hu[i];
        dv[i] = P.hv[i];
        di[i] = P.hins[i];
    }
}

// The counters' state at a batch start: every count zero, no conflict.
__device__ __forceinline__ u64 counter_reset_value(int64_t i) {
    return (i == C_CONF0 || i == C_CONF1) ? ~0ull : 0ull;
}

// One workgroup of NT threads (a constant, as in grid_coords).
template <int NT>
__device__ __forceinline__ GC block_coords() {
    GC c;
    c.tid = threadIdx.x;
    c.nthreads = NT;
    c.wave = c.tid / MF_WAVE;
    c.nwaves = c.nthreads / MF_WAVE;
    c.lane = lane_id();
    return c;
}

#define MF_KERNEL(name, body)                                                   \
    __global__ __launch_bounds__(256) void name(BP P) {                        \
        GC c = grid_coords();                                                   \
        body;                                                                   \
    }

MF_KERNEL(k_init, ph_init(P, c))
MF_KERNEL(k_claim, if (!P.hu && c.tid == 0) P.ctr[MF_CS * (C_T0)] = wall_clock64(); ph_claim(P, c))
MF_KERNEL(k_resolve, ph_resolve(P, c))
MF_KERNEL(k_append, ph_append(P, c))
MF_KERNEL(k_partner, ph_partner(P, c))
MF_KERNEL(k_kcount, ph_kcount(P, c))
MF_KERNEL(k_pu_scan, ph_pu_scan(P, c))
MF_KERNEL(k_fa_count, ph_fa_count(P, c))
MF_KERNEL(k_fa_alloc, ph_fa_alloc(P, c))
MF_KERNEL(k_fa_fill, ph_fa_fill(P, c))
MF_KERNEL(k_comp_collect, ph_comp_collect(P, c))
MF_KERNEL(k_comp_decide, ph_comp_decide(P, c))
MF_KERNEL(k_stats, ph_stats(P, c))

// The round kernels come in workgroups of NT = 128, 256 or 512 threads
// (a tuning knob); they work per wave, so only the grid's shape changes.
template <int NT>
__device__ __forceinline__ GC grid_coords_nt() {
    GC c;
    c.tid = (int64_t)blockIdx.x * NT + threadIdx.x;
    c.nthreads = (int64_t)gridDim.x * NT;
    c.wave = c.tid / MF_WAVE;
    c.nwaves = c.nthreads / MF_WAVE;
    c.lane = lane_id();
    return c;
}

// par bit 1: the batch's first round, which stamps the end of the prologue
// (no commit staged its edges).
template <int NT>
__global__ __launch_bounds__(NT) void k_spec(BP P, int par) {
    __shared__ int cur[NT / MF_WAVE];
    GC c = grid_coords_nt<NT>();
    const bool first = (par & 2) != 0;
    const int4 st = !first && c.wave < P.nstage ? P.stage[c.wave] : make_int4(-1, -1, 0, -1);
    if (first) {
        if (c.tid == 0) P.ctr[MF_CS * (C_TP)] = wall_clock64();
        par &= 1;
    }
    RoundView v = round_view(P, par, P.window);
    if (v.live) ph_spec_w(P, c, v, cur, st);
}
template <int NT>
__global__ __launch_bounds__(NT) void k_validate(BP P, int par) {
    GC c = grid_coords_nt<NT>();
    // the segment's first line (every wave of the round grid has one)
    const uint4 r0 = c.lane < MF_RLOG_LINE ? P.rlog[c.wave * MF_RLOG + c.lane] : make_uint4(0u, 0u, 0u, 0u);
    RoundView v = round_view(P, par, P.window);
    if (v.live) ph_validate_w(P, c, v, par, r0);
}
template <int NT>
__global__ __launch_bounds__(NT) void k_commit(BP P, int par) {
    GC c = grid_coords_nt<NT>();
    RoundView v = round_view(P, par, P.window);
    if (v.live) ph_commit_w(P, c, v, par);
    else if (c.tid == 0) advance_round(P, par, INT64_MAX, P.window);
}
__global__ __launch_bounds__(256) void k_comp_apply(BP P, int par) {
    GC c = grid_coords();
    ph_comp_apply(P, c, par);
}

// Barrier between phases inside one workgroup.  Atomics are performed in L2
// while a plain load may be served by the CU's L1, and a bare workgroup
// barrier neither waits for outstanding vector memory operations nor drops
// L1.  All waves of the workgroup sit on one CU and share one L2, so waiting
// for the wave's own memory operations before the barrier and invalidating
// the CU's L1 after it makes every phase see the previous phase's atomics
// and stores; the L2 itself needs no write-back (an agent-scope fence would
// write back and invalidate the whole L2 at every phase).
__device__ __forceinline__ void phase_barrier() {
    __builtin_amdgcn_s_waitcnt(0);                // own stores and atomics done
    __syncthreads();
    asm volatile("buffer_inv sc0" ::: "memory");  // drop this CU's L1 lines
}

// Stamps the end of the batch on the device clock, then copies the counters
// and the pool top to the pinned host block of the batch (one workgroup), so
// the host reads the batch's results without a copy on the stream.  At the
// batch's end (`last`) each counter is reset for the next batch right after
// it is copied; a halted tail keeps them for the rounds queued after it.
__device__ __forceinline__ void publish(const BP& P, const GC& c, bool last) {
    if (c.tid == 0) P.ctr[MF_CS * (C_T1)] = wall_clock64();
    phase_barrier();
    for (int64_t i = c.tid; i < C_NCOUNTERS; i += c.nthreads) {
        P.h_out[i] = P.ctr[MF_CS * (i)];
        if (last) P.ctr[MF_CS * (i)] = counter_reset_value(i);
    }
    if (c.tid == 0) P.h_out[C_NCOUNTERS] = *P.g.pool_top;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
}

// Remaining speculative rounds in one workgroup; returns false when a loop
// bound was hit.  `ran`: some round ran (nothing before the first needs a
// barrier: the kernel's start sees what earlier kernels wrote).
__device__ __forceinline__ bool finish_rounds(const BP& P, const GC& c, int par, bool& ran) {
    uint32_t guard = 0;
    ran = false;
    while (true) {
        if (ran) phase_barrier();
        const int64_t fw = P.fin_window > 0 ? P.fin_window : MF_FIN_WINDOW;
        const int64_t win = P.window < fw ? P.window : fw;
        RoundView v = round_view(P, par, win);
        if (!v.live) return true;
        ran = true;
        if ((int64_t)guard++ > P.U + 1) {  // every round commits its first edge
            if (c.tid == 0) P.ctr[MF_CS * (C_GUARD)] = 1;
            return false;
        }
        ph_classify_l(P, c, v, par);
        phase_barrier();
        ph_spec_l(P, c, par, v.R);
        phase_barrier();
        ph_validate_l(P, c, v, par);
        phase_barrier();
        ph_commit_l(P, c, par);
        if (c.tid == 0) advance_round(P, par, round_conflict(P, par), win);
        par ^= 1;
    }
}

// Completion iterations over the collected candidates in one workgroup.
__device__ __forceinline__ void finish_completion(const BP& P, const GC& c) {
    if (!P.completion) return;
    for (uint32_t k = 0;; k++) {
        phase_barrier();
        int cp = (int)(k & 1u);
        ph_comp_decide(P, c);
        phase_barrier();
        ph_comp_apply(P, c, cp);
        phase_barrier();
        if (P.ctr[MF_CS * (C_UNDEC0 + cp)] == 0) break;
        if ((u64)k > P.ctr[MF_CS * (C_CAND)] + 1) {  // the smallest open candidate decides each time
            if (c.tid == 0) P.ctr[MF_CS * (C_GUARD)] = 2;
            return;
        }
    }
}

// The rounds a tail still finds (parity par first), the rule counters, the
// completion and the publication, in one workgroup; the end of the rounds
// and of the counters are stamped on the device clock.
__device__ __forceinline__ void tail_phases(const BP& P, const GC& c, int par) {
    bool ran = false;
    if (finish_rounds(P, c, par, ran)) {
        if (ran) phase_barrier();
        if (c.tid == 0) P.ctr[MF_CS * (C_TR)] = wall_clock64();
        ph_stats(P, c);
        if (P.completion) {
            ph_comp_collect(P, c);
            if (c.tid == 0) P.ctr[MF_CS * (C_TS)] = wall_clock64();
            finish_completion(P, c);
        } else if (c.tid == 0) {
            P.ctr[MF_CS * (C_TS)] = wall_clock64();
        }
    }
    publish(P, c, true);
}

// The tail of a batch in one workgroup, after the queued grid-wide rounds:
// whatever rounds they left, the rule counters, the completion candidates
// and iterations, then the counters are published to the host.  Its phases
// are separated by workgroup barriers, far cheaper than grid-wide
// synchronisation on a multi-die GPU, and each of them has few items.  A
// tail that may halt (P.may_halt: the host queued the rounds without a
// reliable estimate and nothing of this graph behind them) stops instead
// when more than MF_HALT_EDGES edges are left, publishing C_HALT = 1, and
// the host queues more grid-wide rounds and another tail.
__global__ __launch_bounds__(1024) void k_finish(BP P, int par) {
    GC c = block_coords<1024>();
    if (P.may_halt) {
        phase_barrier();
        const int64_t start = (int64_t)P.ctr[MF_CS * (C_START0 + par)];
        const bool halt = P.U - start > MF_HALT_EDGES;
        if (c.tid == 0) P.ctr[MF_CS * (C_HALT)] = halt ? 1ull : 0ull;
        if (halt) {
            publish(P, c, false);
            return;
        }
    }
    tail_phases(P, c, par);
}

__global__ __launch_bounds__(256) void k_publish(BP P) {
    GC c = block_coords<256>();
    publish(P, c, true);
}

// ---------------------------------------------------------------- serving kernel
// A batch in one workgroup: (whole) the prologue's phases, then the
// speculative rounds (window P.fin_window), the rule counters, the
// completion and the publication, separated by workgroup barriers instead
// of kernel boundaries.  Every phase is the grid-stride body the grid-wide
// kernels run, here over the workgroup's threads.  Without `whole` the
// prologue ran as grid-wide kernels before the item was published.
__device__ __forceinline__ void batch_in_workgroup(const BP& P, const GC& c, bool whole) {
    if (!whole) {
        if (c.tid == 0) P.ctr[MF_CS * (C_TP)] = wall_clock64();
        tail_phases(P, c, 0);
        return;
    }
    if (P.hu) ph_init(P, c);
    else if (c.tid == 0) P.ctr[MF_CS * (C_T0)] = wall_clock64();
    phase_barrier();
    ph_claim(P, c);
    phase_barrier();
    ph_resolve(P, c);
    phase_barrier();
    ph_append(P, c);
    phase_barrier();
    ph_partner(P, c);
    phase_barrier();
    ph_kcount(P, c);
    phase_barrier();
    ph_pu_scan(P, c);
    phase_barrier();
    ph_fa_count(P, c);
    phase_barrier();
    ph_fa_alloc(P, c);
    phase_barrier();
    ph_fa_fill(P, c);
    phase_barrier();
    if (c.tid == 0) P.ctr[MF_CS * (C_TP)] = wall_clock64();
    tail_phases(P, c, 0);
}

__device__ __forceinline__ u64 host_word(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The claim of the next item by the workgroup's first thread: the item's
// index, or -1 when the host asked the kernel to stop or no item came for
// idle_ticks of the device clock.  Waits sleep about 3 us between looks.
__device__ int64_t serve_claim(ServeQueue* q, u64* head, u64 idle_ticks) {
    const u64 t0 = wall_clock64();
    while (true) {
        if (host_word(&q->stop)) return -1;
        const u64 tail = host_word(&q->tail);
        const u64 h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h < tail) {
            if (atomicCAS(head, h, h + 1) == h) return (int64_t)h;
            continue;  // another workgroup took it
        }
        if (wall_clock64() - t0 > idle_ticks) return -1;
        __builtin_amdgcn_s_sleep(127);
    }
}

// Workgroups take whole batches from the queue, any graph's, one at a time.
// A batch may run on a different die than its graph's last one, or after
// kernels of its graph's own stream: each item starts with an agent-scope
// acquire (this die's L2 drops lines others may have written) and its
// publication ends with a system-scope release (the batch's writes leave
// the L2), so every batch sees its graph's state as the last one left it.
template <int NT>
__global__ __launch_bounds__(NT) void k_serve(ServeQueue* q, u64* head, u64 idle_ticks) {
    __shared__ BP sP;
    __shared__ int64_t s_item;
    __shared__ int s_whole;
    GC c = block_coords<NT>();
    while (true) {
        if (c.tid == 0) s_item = serve_claim(q, head, idle_ticks);
        __syncthreads();
        const int64_t item = s_item;
        if (item < 0) return;  // every thread read the same value
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&q->items[item % MF_SERVE_CAP].P);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&sP);
        for (int i = (int)c.tid; i < (int)(sizeof(BP) / 4); i += (int)c.nthreads)
            dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (c.tid == 0)
            s_whole = __hip_atomic_load(&q->items[item % MF_SERVE_CAP].whole, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_SYSTEM) != 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        batch_in_workgroup(sP, c, s_whole != 0);
        __syncthreads();
        if (c.tid == 0)
            __hip_atomic_store(&sP.h_out[MF_DONE_WORD], (u64)item + 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();  // sP is rewritten by the next item only after this
    }
}

// ---------------------------------------------------------------- launchers
// Grids are sized to what the GPU holds at once: a kernel's workgroups beyond
// its occupancy wait for earlier ones to finish, and in these latency-bound
// phases that wait is a whole evaluation chain (measured 10-17 us per round
// at 8 workgroups per CU when registers allow 7).  The occupancy of each
// kernel is queried once per process and device.
namespace {
struct OccEntry {
    const void* fn;
    int device;
    int per_cu;
};
std::mutex g_occ_mu;
OccEntry g_occ[64];
int g_nocc = 0;

int occ_per_cu(const void* fn, int threads) {
    int dev = 0;
    hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_occ_mu);
    for (int i = 0; i < g_nocc; i++)
        if (g_occ[i].fn == fn && g_occ[i].device == dev) return g_occ[i].per_cu;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, 0) != hipSuccess || nb < 1) nb = 1;
    if (g_nocc < 64) g_occ[g_nocc++] = OccEntry{fn, dev, nb};
    return nb;
}

// L.blocks is the cap (8 workgroups of 256 threads per CU); the grid is
// min(cap, resident), both counted in workgroups of `threads`.
int grid_for(const void* fn, const LaunchCfg& L, int threads = 256) {
    int per_cu = occ_per_cu(fn, threads);
    int resident = per_cu * (L.blocks / 8);
    const int cap = (int)((int64_t)L.blocks * 256 / threads);
    return resident < cap ? resident : cap;
}

// The round kernels of one workgroup size.
struct RoundFns {
    const void* spec;
    const void* validate;
    const void* commit;
};
RoundFns round_fns(int threads) {
    if (threads == 128) return {(const void*)k_spec<128>, (const void*)k_validate<128>, (const void*)k_commit<128>};
    if (threads == 512) return {(const void*)k_spec<512>, (const void*)k_validate<512>, (const void*)k_commit<512>};
    return {(const void*)k_spec<256>, (const void*)k_validate<256>, (const void*)k_commit<256>};
}
}  // namespace

enum GridClass { G_FULL, G_UPD, G_CLAIM, G_WAVE_UPD, G_SCAN, G_SCAN32 };

// The grid of a phase for this batch: what its work needs (see LaunchCfg),
// at least 8 workgroups, at most what the GPU holds at once.
int grid_cls(const void* fn, const LaunchCfg& L, GridClass cls) {
    int full = grid_for(fn, L);
    if (L.U < 0 || cls == G_FULL) return full;
    int64_t want = cls == G_UPD        ? (L.U + 255) / 256
                   : cls == G_CLAIM    ? (claim_threads(L.U) + 255) / 256
                   : cls == G_WAVE_UPD ? (L.U + 3) / 4
                   : cls == G_SCAN     ? (L.n + 1023) / 1024
                                       : ((L.n + 31) / 32 + 255) / 256;
    if (want < 8) want = 8;
    return want < full ? (int)want : full;
}

// Speculation and validation share one grid: wave w of the speculation
// writes read-log segment w, which wave w's slots of the validation check.
// L.round_grid > 0 sets it for the batch; otherwise L.round_per_cu > 0 sets
// it (workgroups per CU) instead of the occupancy.
// The grid never holds more waves than 4 per workgroup of L.blocks: the
// read log has one segment per such wave.
int mf_round_grid(const LaunchCfg& L) {
    if (L.round_grid > 0) return L.round_grid;
    const int cap = (int)((int64_t)L.blocks * 256 / L.round_threads);
    if (L.round_per_cu > 0) {
        int g = L.round_per_cu * (L.blocks / 8);
        return g < cap ? g : cap;
    }
    const RoundFns f = round_fns(L.round_threads);
    int a = grid_for(f.spec, L, L.round_threads), b = grid_for(f.validate, L, L.round_threads);
    return a < b ? a : b;
}

#define LAUNCH(k, ...) \
    hipLaunchKernelGGL(k, dim3(grid_for((const void*)k, L)), dim3(256), 0, L.stream, __VA_ARGS__)
#define LAUNCH_C(k, cls, ...) \
    hipLaunchKernelGGL(k, dim3(grid_cls((const void*)k, L, cls)), dim3(256), 0, L.stream, __VA_ARGS__)
#define LAUNCH_ROUND(k, NT, ...) \
    hipLaunchKernelGGL(k<NT>, dim3(mf_round_grid(L)), dim3(NT), 0, L.stream, __VA_ARGS__)

void mf_launch_batch_prologue(const BP& P, const LaunchCfg& L) {
    if (P.hu) LAUNCH_C(k_init, G_UPD, P);
    LAUNCH_C(k_claim, G_CLAIM, P);
    LAUNCH_C(k_resolve, G_UPD, P);
    LAUNCH_C(k_append, G_WAVE_UPD, P);
    LAUNCH_C(k_partner, G_UPD, P);
    LAUNCH_C(k_kcount, G_WAVE_UPD, P);
    LAUNCH_C(k_pu_scan, P.pu_list ? G_SCAN32 : G_SCAN, P);
    LAUNCH_C(k_fa_count, G_WAVE_UPD, P);
    LAUNCH_C(k_fa_alloc, G_UPD, P);
    LAUNCH_C(k_fa_fill, G_WAVE_UPD, P);
}

template <int NT>
static void launch_round_nt(const BP& P, const LaunchCfg& L, int par, bool first) {
    LAUNCH_ROUND(k_spec, NT, P, par | (first ? 2 : 0));
    LAUNCH_ROUND(k_validate, NT, P, par);
    if (L.round_grid > 0) LAUNCH_ROUND(k_commit, NT, P, par);
    else
        hipLaunchKernelGGL(k_commit<NT>, dim3(grid_for((const void*)k_commit<NT>, L, NT)), dim3(NT), 0, L.stream, P,
                           par);
}

void mf_launch_round(const BP& P, const LaunchCfg& L, int par, bool first) {
    if (L.round_threads == 128) launch_round_nt<128>(P, L, par, first);
    else if (L.round_threads == 512) launch_round_nt<512>(P, L, par, first);
    else launch_round_nt<256>(P, L, par, first);
}

void mf_launch_finish(const BP& P, const LaunchCfg& L, int par) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, L.stream, P, par);
}

void mf_launch_completion_collect(const BP& P, const LaunchCfg& L) { LAUNCH(k_comp_collect, P); }

void mf_launch_stats(const BP& P, const LaunchCfg& L) { LAUNCH(k_stats, P); }

void mf_launch_publish(const BP& P, const LaunchCfg& L) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(256), 0, L.stream, P);
}

void mf_launch_serve(ServeQueue* q, u64* head, int grid, int threads, u64 idle_ticks, hipStream_t st) {
    if (threads >= 1024) hipLaunchKernelGGL(k_serve<1024>, dim3(grid), dim3(1024), 0, st, q, head, idle_ticks);
    else if (threads >= 512) hipLaunchKernelGGL(k_serve<512>, dim3(grid), dim3(512), 0, st, q, head, idle_ticks);
    else hipLaunchKernelGGL(k_serve<256>, dim3(grid), dim3(256), 0, st, q, head, idle_ticks);
}

void mf_launch_completion_round(const BP& P, const LaunchCfg& L, int par) {
    LAUNCH(k_comp_decide, P);
    LAUNCH(k_comp_apply, P, par);
}
