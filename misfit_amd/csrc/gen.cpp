// gen.cpp -- seeded synthetic graphs, initial MIS and update streams.
//
// Randomness: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) streams; bounded
// integers by Lemire's multiply-shift with rejection (ACM TOMACS 2019).
// Parallel generators split the work into fixed chunks, each with its own
// stream seeded from (seed, chunk), so results do not depend on the thread
// count.  Every graph is returned as a symmetric CSR without self-loops or
// duplicates, rows ascending.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "../../include/misfit.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // uniform in [0, bound)
    uint64_t below(uint64_t bound) {
        __uint128_t m = (__uint128_t)next() * bound;
        uint64_t l = (uint64_t)m;
        if (l < bound) {
            uint64_t t = (0 - bound) % bound;
            while (l < t) {
                m = (__uint128_t)next() * bound;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t stream_seed(uint64_t seed, uint64_t chunk) {
    Rng r(seed ^ (0x632be59bd9b4e019ull * (chunk + 1)));
    return r.next();
}

// Directed pair list (both directions already present) -> CSR, rows sorted and
// deduplicated, self-loops dropped.
int build_csr(int64_t n, std::vector<uint64_t>& pairs, int64_t* nindex, int32_t* nlist, int64_t* m_out) {
    std::vector<int64_t> deg((size_t)n + 1, 0);
    for (uint64_t p : pairs) deg[(size_t)(p >> 32)]++;
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + deg[i];
    std::vector<int32_t> tmp((size_t)off[n]);
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (uint64_t p : pairs) tmp[(size_t)cur[(size_t)(p >> 32)]++] = (int32_t)(p & 0xFFFFFFFFu);
    pairs.clear();
    pairs.shrink_to_fit();
    std::vector<int64_t> kept((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t x = 0; x < n; x++) {
        int32_t* b = tmp.data() + off[x];
        int32_t* e = tmp.data() + off[x + 1];
        std::sort(b, e);
        int32_t* w = b;
        for (int32_t* r = b; r < e; r++) {
            if (*r == (int32_t)x) continue;
            if (w > b && *(w - 1) == *r) continue;
            *w++ = *r;
        }
        kept[x] = w - b;
    }
    int64_t m = 0;
    for (int64_t x = 0; x < n; x++) m += kept[x];
    *m_out = m;
    if (!nindex) return MISFIT_OK;
    nindex[0] = 0;
    for (int64_t x = 0; x < n; x++) nindex[x + 1] = nindex[x] + kept[x];
    if (nlist) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t x = 0; x < n; x++)
            memcpy(nlist + nindex[x], tmp.data() + off[x], (size_t)kept[x] * 4);
    }
    return MISFIT_OK;
}

inline uint64_t dpair(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

// CSR of the directed pairs a chunked generator emits, without holding the
// pair list: gen(ch, emit) calls emit(src, dst) for every pair of chunk ch
// and must emit the same pairs each time it is called (each chunk draws from
// its own seeded stream).  Pass 1 counts the row lengths, pass 2 scatters
// the entries to their rows through per-row cursors; the order inside a row
// then depends on the threads, but every row is sorted, deduplicated and
// stripped of self-loops, so the result does not.
template <class Gen>
int build_csr_chunked(int64_t n, int64_t nch, const Gen& gen, int64_t* nindex, int32_t* nlist, int64_t* m_out) {
    std::vector<int64_t> off((size_t)n + 1, 0);
    int64_t* offp = off.data();
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ch = 0; ch < nch; ch++)
        gen(ch, [&](uint32_t a, uint32_t) { __atomic_fetch_add(&offp[(size_t)a + 1], (int64_t)1, __ATOMIC_RELAXED); });
    for (int64_t i = 0; i < n; i++) off[i + 1] += off[i];
    std::vector<int32_t> tmp((size_t)off[n]);
    int32_t* tp = tmp.data();
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        int64_t* cp = cur.data();
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t ch = 0; ch < nch; ch++)
            gen(ch, [&](uint32_t a, uint32_t b) {
                tp[__atomic_fetch_add(&cp[a], (int64_t)1, __ATOMIC_RELAXED)] = (int32_t)b;
            });
    }
    std::vector<int64_t> kept((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t x = 0; x < n; x++) {
        int32_t* b = tp + off[x];
        int32_t* e = tp + off[x + 1];
        std::sort(b, e);
        int32_t* w = b;
        for (int32_t* r = b; r < e; r++) {
            if (*r == (int32_t)x) continue;
            if (w > b && *(w - 1) == *r) continue;
            *w++ = *r;
        }
        kept[x] = w - b;
    }
    int64_t m = 0;
    for (int64_t x = 0; x < n; x++) m += kept[x];
    *m_out = m;
    if (!nindex) return MISFIT_OK;
    nindex[0] = 0;
    for (int64_t x = 0; x < n; x++) nindex[x + 1] = nindex[x] + kept[x];
    if (nlist) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t x = 0; x < n; x++) memcpy(nlist + nindex[x], tp + off[x], (size_t)kept[x] * 4);
    }
    return MISFIT_OK;
}

// Row of CSR entry p (the largest x with nindex[x] <= p), through a coarse
// table of the row holding every 4,096th entry, so a lookup searches only
// the rows between two table entries instead of all n.
struct EntryRows {
    const int64_t* nindex;
    int32_t n;
    std::vector<int32_t> at;  // at[j] = row of entry j * 4096
    EntryRows(const int64_t* ni, int32_t nn) : nindex(ni), n(nn) {
        const int64_t m = ni[nn];
        at.resize((size_t)(m >> 12) + 2);
        int32_t x = 0;
        for (size_t j = 0; j < at.size(); j++) {
            const int64_t p = std::min<int64_t>((int64_t)j << 12, m);
            while (x + 1 <= nn && ni[x + 1] <= p) x++;
            at[j] = x;
        }
    }
    int32_t row(int64_t p) const {
        const size_t j = (size_t)(p >> 12);
        const int32_t lo = at[j], hi = std::min<int32_t>(at[j + 1] + 1, n);
        return (int32_t)(std::upper_bound(nindex + lo, nindex + hi + 1, p) - nindex - 1);
    }
};

bool base_adjacent(const int64_t* nindex, const int32_t* nlist, int32_t u, int32_t v) {
    const int32_t* b = nlist + nindex[u];
    const int32_t* e = nlist + nindex[u + 1];
    return std::binary_search(b, e, v);
}

}  // namespace

extern "C" int misfit_gen_er(int32_t n, int64_t M, uint64_t seed, int64_t* nindex, int32_t* nlist, int64_t* m) {
    if (n < 2 || M < 0 || !m) return MISFIT_E_ARG;
    if ((double)M > 0.5 * (double)n * (n - 1)) return MISFIT_E_ARG;
    Rng r(seed);
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)M * 2);
    std::vector<uint64_t> pairs;
    pairs.reserve((size_t)M * 2);
    while ((int64_t)seen.size() < M) {
        uint32_t a = (uint32_t)r.below((uint64_t)n), b = (uint32_t)r.below((uint64_t)n);
        if (a == b) continue;
        uint32_t lo = std::min(a, b), hi = std::max(a, b);
        if (!seen.insert(dpair(lo, hi)).second) continue;
        pairs.push_back(dpair(lo, hi));
        pairs.push_back(dpair(hi, lo));
    }
    return build_csr(n, pairs, nindex, nlist, m);
}

extern "C" int misfit_gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                               int64_t* nindex, int32_t* nlist, int64_t* m) {
    if (scale < 1 || scale > 30 || edgefactor < 1 || !m) return MISFIT_E_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return MISFIT_E_ARG;
    const int64_t n = (int64_t)1 << scale;
    const int64_t E = (int64_t)edgefactor * n;
    const int64_t CH = 1 << 16;
    const int64_t nch = (E + CH - 1) / CH;
    const double ab = a + b, abc = a + b + c;
    // chunk ch draws edges [ch*CH, (ch+1)*CH) from its own stream; each
    // edge is emitted in both directions
    auto gen = [&](int64_t ch, auto&& emit) {
        Rng r(stream_seed(seed, (uint64_t)ch));
        int64_t e0 = ch * CH, e1 = std::min(E, e0 + CH);
        for (int64_t e = e0; e < e1; e++) {
            uint32_t u = 0, v = 0;
            for (int l = 0; l < scale; l++) {
                double q = r.unit();
                uint32_t bu = 0, bv = 0;
                if (q < a) {
                } else if (q < ab) {
                    bv = 1;
                } else if (q < abc) {
                    bu = 1;
                } else {
                    bu = 1;
                    bv = 1;
                }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            emit(u, v);
            emit(v, u);
        }
    };
    return build_csr_chunked(n, nch, gen, nindex, nlist, m);
}

extern "C" int misfit_gen_rgg(int32_t n, double avg_degree, uint64_t seed, int64_t* nindex, int32_t* nlist,
                              int64_t* m) {
    if (n < 2 || avg_degree <= 0 || !m) return MISFIT_E_ARG;
    const double PI = 3.14159265358979323846;
    const double r = std::sqrt(avg_degree / (PI * (double)n));
    const double r2 = r * r;
    int64_t G = (int64_t)std::floor(1.0 / r);
    if (G < 1) G = 1;
    if (G > 65536) G = 65536;
    const double cell = 1.0 / (double)G;
    std::vector<double> xs((size_t)n), ys((size_t)n);
    const int64_t CH = 1 << 16;
    const int64_t nch = ((int64_t)n + CH - 1) / CH;
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < nch; ch++) {
        Rng rr(stream_seed(seed, (uint64_t)ch));
        for (int64_t i = ch * CH; i < std::min<int64_t>(n, (ch + 1) * CH); i++) {
            xs[i] = rr.unit();
            ys[i] = rr.unit();
        }
    }
    auto cell_of = [&](int64_t i) {
        int64_t cx = std::min<int64_t>(G - 1, (int64_t)(xs[i] / cell));
        int64_t cy = std::min<int64_t>(G - 1, (int64_t)(ys[i] / cell));
        return cy * G + cx;
    };
    std::vector<int64_t> cstart((size_t)(G * G + 1), 0);
    std::vector<int64_t> pc((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        pc[i] = cell_of(i);
        cstart[(size_t)pc[i] + 1]++;
    }
    for (int64_t k = 0; k < G * G; k++) cstart[k + 1] += cstart[k];
    std::vector<int32_t> bucket((size_t)n);
    {
        std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
        for (int64_t i = 0; i < n; i++) bucket[(size_t)cur[(size_t)pc[i]]++] = (int32_t)i;
    }
    std::vector<int64_t> deg((size_t)n, 0);
    auto for_nb = [&](int64_t i, auto&& fn) {
        int64_t c0 = pc[i], cx = c0 % G, cy = c0 / G;
        for (int64_t dy = -1; dy <= 1; dy++) {
            int64_t yy = cy + dy;
            if (yy < 0 || yy >= G) continue;
            for (int64_t dx = -1; dx <= 1; dx++) {
                int64_t xx = cx + dx;
                if (xx < 0 || xx >= G) continue;
                int64_t k = yy * G + xx;
                for (int64_t t = cstart[k]; t < cstart[k + 1]; t++) {
                    int32_t j = bucket[t];
                    if (j == (int32_t)i) continue;
                    double ddx = xs[i] - xs[j], ddy = ys[i] - ys[j];
                    if (ddx * ddx + ddy * ddy < r2) fn(j);
                }
            }
        }
    };
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; i++) {
        int64_t d = 0;
        for_nb(i, [&](int32_t) { d++; });
        deg[i] = d;
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) total += deg[i];
    *m = total;
    if (!nindex) return MISFIT_OK;
    nindex[0] = 0;
    for (int64_t i = 0; i < n; i++) nindex[i + 1] = nindex[i] + deg[i];
    if (!nlist) return MISFIT_OK;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; i++) {
        int32_t* w = nlist + nindex[i];
        for_nb(i, [&](int32_t j) { *w++ = j; });
        std::sort(nlist + nindex[i], nlist + nindex[i + 1]);
    }
    return MISFIT_OK;
}

// tools/compute_initial_mis.py:32-43: ascending id, a vertex joins unless a
// member already blocked it.
extern "C" int misfit_greedy_mis(int32_t n, const int64_t* nindex, const int32_t* nlist, uint8_t* mis) {
    if (n < 1 || !nindex || !mis) return MISFIT_E_ARG;
    std::vector<uint8_t> blocked((size_t)n, 0);
    for (int32_t v = 0; v < n; v++) {
        mis[v] = 0;
        if (blocked[v]) continue;
        mis[v] = 1;
        for (int64_t p = nindex[v]; p < nindex[v + 1]; p++) blocked[(size_t)nlist[p]] = 1;
    }
    return MISFIT_OK;
}

extern "C" int misfit_gen_updates(int kind, int32_t n, const int64_t* nindex, const int32_t* nlist, int64_t count,
                                  uint64_t seed, int32_t* u, int32_t* v) {
    if (n < 2 || count < 0 || !u || !v || !nindex) return MISFIT_E_ARG;
    // a stream of its own, so a graph and its updates drawn with the same
    // seed are independent
    Rng r(stream_seed(seed, 0x75706461746573ull));
    const int64_t m = nindex[n];
    if (kind == 0) {
        for (int64_t i = 0; i < count; i++) {
            int32_t a, b;
            do {
                a = (int32_t)r.below((uint64_t)n);
                b = (int32_t)r.below((uint64_t)n);
            } while (a == b);
            u[i] = a;
            v[i] = b;
        }
        return MISFIT_OK;
    }
    if (kind == 1) {
        int64_t ndel = count / 2, nins = count - ndel;
        if (ndel > m / 2) return MISFIT_E_ARG;
        const EntryRows rows(nindex, n);
        std::unordered_set<uint64_t> used;
        used.reserve((size_t)count * 2);
        std::vector<uint64_t> out;
        out.reserve((size_t)count);
        while ((int64_t)out.size() < ndel) {
            int64_t p = (int64_t)r.below((uint64_t)m);
            int32_t src = rows.row(p);
            int32_t dst = nlist[p];
            if (src >= dst) continue;  // each undirected edge once: the entry with src < dst
            if (!used.insert(dpair((uint32_t)src, (uint32_t)dst)).second) continue;
            out.push_back(dpair((uint32_t)src, (uint32_t)dst));
        }
        while ((int64_t)out.size() < ndel + nins) {
            int32_t a = (int32_t)r.below((uint64_t)n), b = (int32_t)r.below((uint64_t)n);
            if (a == b) continue;
            int32_t lo = std::min(a, b), hi = std::max(a, b);
            if (base_adjacent(nindex, nlist, lo, hi)) continue;
            if (!used.insert(dpair((uint32_t)lo, (uint32_t)hi)).second) continue;
            out.push_back(dpair((uint32_t)lo, (uint32_t)hi));
        }
        for (int64_t i = (int64_t)out.size() - 1; i > 0; i--) {
            int64_t j = (int64_t)r.below((uint64_t)(i + 1));
            std::swap(out[(size_t)i], out[(size_t)j]);
        }
        for (int64_t i = 0; i < count; i++) {
            int32_t a = (int32_t)(out[(size_t)i] >> 32), b = (int32_t)(out[(size_t)i] & 0xFFFFFFFFu);
            if (r.next() & 1) std::swap(a, b);
            u[i] = a;
            v[i] = b;
        }
        return MISFIT_OK;
    }
    if (kind == 2) {
        std::vector<int64_t> cum((size_t)n + 1, 0);
        for (int32_t x = 0; x < n; x++) cum[x + 1] = cum[x] + (nindex[x + 1] - nindex[x]) + 1;
        const int64_t tot = cum[n];
        auto draw = [&]() {
            int64_t t = (int64_t)r.below((uint64_t)tot);
            return (int32_t)(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin() - 1);
        };
        for (int64_t i = 0; i < count; i++) {
            int32_t a, b;
            do {
                a = draw();
                b = draw();
            } while (a == b);
            u[i] = a;
            v[i] = b;
        }
        return MISFIT_OK;
    }
    if (kind >= 4 && kind <= 6) {
        // Distinct existing undirected edges as (lo, hi), shuffled; all
        // deletions against the loaded graph.  kind 4: a uniform CSR entry
        // (tools/generate_test_batches.py:55-78); kind 5: a vertex drawn
        // proportionally to 1/degree, then a uniform neighbour
        // (tools/generate_low_degree_batches.py:56-100); kind 6: a vertex
        // drawn proportionally to its degree, then a uniform neighbour
        // (tools/generate_preferential_batches.py:56-98).  Fails, like those
        // scripts, when the graph cannot supply count distinct edges.
        if (count > m) return MISFIT_E_ARG;
        std::vector<double> cumw;
        if (kind == 5) {
            cumw.assign((size_t)n + 1, 0.0);
            for (int32_t x = 0; x < n; x++) {
                int64_t d = nindex[x + 1] - nindex[x];
                cumw[x + 1] = cumw[x] + (d > 0 ? 1.0 / (double)d : 0.0);
            }
            if (cumw[n] <= 0) return MISFIT_E_ARG;
        }
        if (m <= 0) return MISFIT_E_ARG;
        const EntryRows rows(nindex, n);
        std::unordered_set<uint64_t> used;
        used.reserve((size_t)count * 2);
        std::vector<uint64_t> out;
        out.reserve((size_t)count);
        const int64_t max_draws = 200 * count + (1 << 20);
        for (int64_t t = 0; (int64_t)out.size() < count; t++) {
            if (t >= max_draws) return MISFIT_E_ARG;
            int32_t src;
            int64_t p;
            if (kind == 5) {
                double w = r.unit() * cumw[n];
                src = (int32_t)(std::upper_bound(cumw.begin(), cumw.end(), w) - cumw.begin() - 1);
                if (src < 0) src = 0;
                if (src >= n) src = n - 1;
                int64_t d = nindex[src + 1] - nindex[src];
                if (d <= 0) continue;
                p = nindex[src] + (int64_t)r.below((uint64_t)d);
            } else if (kind == 6) {
                int64_t q = (int64_t)r.below((uint64_t)m);  // proportional to degree
                src = rows.row(q);
                int64_t d = nindex[src + 1] - nindex[src];
                p = nindex[src] + (int64_t)r.below((uint64_t)d);
            } else {
                p = (int64_t)r.below((uint64_t)m);
                src = rows.row(p);
            }
            int32_t dst = nlist[p];
            if (dst == src) continue;
            uint32_t lo = (uint32_t)std::min(src, dst), hi = (uint32_t)std::max(src, dst);
            if (!used.insert(dpair(lo, hi)).second) continue;
            out.push_back(dpair(lo, hi));
        }
        for (int64_t i = (int64_t)out.size() - 1; i > 0; i--) {
            int64_t j = (int64_t)r.below((uint64_t)(i + 1));
            std::swap(out[(size_t)i], out[(size_t)j]);
        }
        for (int64_t i = 0; i < count; i++) {
            u[i] = (int32_t)(out[(size_t)i] >> 32);
            v[i] = (int32_t)(out[(size_t)i] & 0xFFFFFFFFu);
        }
        return MISFIT_OK;
    }
    if (kind == 3) {
        std::vector<int32_t> ids((size_t)n);
        for (int32_t x = 0; x < n; x++) ids[x] = x;
        std::stable_sort(ids.begin(), ids.end(), [&](int32_t x, int32_t y) {
            return (nindex[x + 1] - nindex[x]) < (nindex[y + 1] - nindex[y]);
        });
        int64_t q = std::max<int64_t>(2, n / 4);
        for (int64_t i = 0; i < count; i++) {
            int32_t a, b;
            do {
                a = ids[(size_t)r.below((uint64_t)q)];
                b = ids[(size_t)r.below((uint64_t)q)];
            } while (a == b);
            u[i] = a;
            v[i] = b;
        }
        return MISFIT_OK;
    }
    return MISFIT_E_ARG;
}
