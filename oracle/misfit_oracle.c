/*
 * misfit_oracle.c -- CPU restatement of MISFIT's incremental-MIS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU product is
 * compared against, and the timed CPU baseline.  It is never linked into,
 * loaded by, or called from the product library (misfit_amd/).  Only tests/,
 * the smoke entry point and the benchmark's CPU-baseline leg load it.
 *
 * What it restates (reference = anonymous-artifacts/MISFIT, paths relative to
 * the reference root):
 *   - .egr reader, both nindex widths ........ core/FlashGraph.h:108-169
 *   - delta overlay (added lists in insertion order, tombstone sets,
 *     un-tombstone / forget)  ............... core/FlashGraph.h:172-276
 *   - HybridGraph::isAdjacent ............... core/FlashGraph.h:291-301
 *   - batch-file parse + insertion flag ...... core/cli_setup.h:66-82,
 *                                               src/misfit/misfit.cu:812-828
 *   - per-batch overlay update ............... src/misfit/misfit.cu:708-714
 *   - 2-hop rows in BFS discovery order ...... src/misfit/misfit.cu:158-202
 *   - candidate (hop-1) rows ................. src/misfit/misfit.cu:550-558
 *   - repair rule ............................ src/misfit/misfit.cu:277-349
 *   - compaction (base minus tombstones, then
 *     additions in insertion order) .......... core/flash_compactor.h:20-86
 *   - greedy ascending-id MIS ................ tools/compute_initial_mis.py:32-43
 * Canonical choices where the reference is implementation-defined (GPU
 * scheduling order, hash-iteration order, sub-batch capacity splits): the
 * whole batch is one unit, no row truncation, and edges are replayed in batch
 * file order.  A maximality completion pass (greedy ascending id over the
 * vertices a batch can uncover) is this project's addition and can be turned
 * off to get the literal rule alone.
 *
 * Two replay engines implement the same rule:
 *   LITERAL  materialises every row exactly as the reference packs them
 *            (seed rows = [s] ++ N1 ++ N2 in BFS order, candidate rows =
 *            adj(x)) -- OpenMP over seeds as in misfit.cu:532-548 -- and then
 *            replays the rule over the packed rows.
 *   LAZY     evaluates the same predicates on the fly (no packed rows), which
 *            makes larger inputs affordable.  It is pinned to LITERAL by the
 *            tests on random inputs.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <omp.h>

#define OR_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ vectors */
typedef struct {
    int32_t* a;
    int64_t n, cap;
} ivec;

static void iv_reserve(ivec* v, int64_t want) {
    if (v->cap >= want) return;
    int64_t c = v->cap ? v->cap : 4;
    while (c < want) c *= 2;
    v->a = (int32_t*)realloc(v->a, (size_t)c * sizeof(int32_t));
    if (!v->a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    v->cap = c;
}
static void iv_push(ivec* v, int32_t x) { iv_reserve(v, v->n + 1); v->a[v->n++] = x; }
static int64_t iv_find(const ivec* v, int32_t x) {
    for (int64_t i = 0; i < v->n; i++) if (v->a[i] == x) return i;
    return -1;
}
/* order-preserving erase, as std::vector::erase (FlashGraph.h:266) */
static void iv_erase_at(ivec* v, int64_t i) {
    memmove(v->a + i, v->a + i + 1, (size_t)(v->n - i - 1) * sizeof(int32_t));
    v->n--;
}
static void iv_free(ivec* v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }

/* ------------------------------------------------------------------ graph */
typedef struct {
    int32_t n;
    int64_t m;             /* base directed entries */
    int wide;              /* 1 if the file used the int64 (-1 sentinel) form */
    int64_t* nindex;       /* n+1 */
    int32_t* nlist;        /* m */
    ivec* added;           /* FlashGraph.h:174 added_adj, insertion order */
    ivec* removed;         /* FlashGraph.h:175 removed_adj, a set */
    int64_t delta_ops;     /* FlashGraph.h:176 delta_edge_count (operations) */
} or_graph;

static or_graph* g_alloc(int32_t n) {
    or_graph* g = (or_graph*)calloc(1, sizeof(or_graph));
    g->n = n;
    g->added = (ivec*)calloc((size_t)n, sizeof(ivec));
    g->removed = (ivec*)calloc((size_t)n, sizeof(ivec));
    return g;
}

OR_EXPORT void or_free(or_graph* g) {
    if (!g) return;
    for (int32_t i = 0; i < g->n; i++) { iv_free(&g->added[i]); iv_free(&g->removed[i]); }
    free(g->added); free(g->removed); free(g->nindex); free(g->nlist); free(g);
}

OR_EXPORT or_graph* or_from_csr(int32_t n, int64_t m, const int64_t* nindex, const int32_t* nlist) {
    or_graph* g = g_alloc(n);
    g->m = m;
    g->wide = 1;
    g->nindex = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    g->nlist = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
    memcpy(g->nindex, nindex, (size_t)(n + 1) * sizeof(int64_t));
    if (m) memcpy(g->nlist, nlist, (size_t)m * sizeof(int32_t));
    return g;
}

/* .egr parse, FlashGraph.h:108-169: int32 nodes; int32 edges or -1 then
 * int64 edges; nindex of n+1 int32 (narrow) or int64 (wide); int32 nlist. */
OR_EXPORT or_graph* or_load_egr(const char* path, char* err, int errlen) {
    FILE* f = fopen(path, "rb");
    if (!f) { snprintf(err, errlen, "could not open %s", path); return NULL; }
    int32_t nodes = 0, e32 = 0;
    int64_t edges = 0;
    int wide = 0;
    if (fread(&nodes, 4, 1, f) != 1) { snprintf(err, errlen, "failed to read nodes"); fclose(f); return NULL; }
    if (fread(&e32, 4, 1, f) != 1) { snprintf(err, errlen, "failed to read edges"); fclose(f); return NULL; }
    if (e32 == -1) {
        if (fread(&edges, 8, 1, f) != 1) { snprintf(err, errlen, "failed to read 64-bit edge count"); fclose(f); return NULL; }
        wide = 1;
    } else {
        if (nodes < 1 || e32 < 0) { snprintf(err, errlen, "invalid header"); fclose(f); return NULL; }
        edges = e32;
    }
    if (nodes < 0 || edges < 0) { snprintf(err, errlen, "invalid header"); fclose(f); return NULL; }
    or_graph* g = g_alloc(nodes);
    g->m = edges;
    g->wide = wide;
    g->nindex = (int64_t*)malloc((size_t)(nodes + 1) * sizeof(int64_t));
    g->nlist = (int32_t*)malloc((size_t)(edges ? edges : 1) * sizeof(int32_t));
    if (wide) {
        if (fread(g->nindex, 8, (size_t)nodes + 1, f) != (size_t)nodes + 1) goto trunc;
    } else {
        int32_t* tmp = (int32_t*)malloc((size_t)(nodes + 1) * 4);
        if (fread(tmp, 4, (size_t)nodes + 1, f) != (size_t)nodes + 1) { free(tmp); goto trunc; }
        for (int64_t i = 0; i <= nodes; i++) g->nindex[i] = tmp[i];
        free(tmp);
    }
    if (edges && fread(g->nlist, 4, (size_t)edges, f) != (size_t)edges) goto trunc;
    fclose(f);
    return g;
trunc:
    snprintf(err, errlen, "truncated file %s", path);
    fclose(f);
    or_free(g);
    return NULL;
}

OR_EXPORT int32_t or_nodes(const or_graph* g) { return g->n; }
OR_EXPORT int64_t or_base_edges(const or_graph* g) { return g->m; }
OR_EXPORT int or_is_wide(const or_graph* g) { return g->wide; }
OR_EXPORT int64_t or_delta_ops(const or_graph* g) { return g->delta_ops; }
OR_EXPORT void or_base_csr(const or_graph* g, int64_t* nindex, int32_t* nlist) {
    memcpy(nindex, g->nindex, (size_t)(g->n + 1) * 8);
    if (g->m) memcpy(nlist, g->nlist, (size_t)g->m * 4);
}

/* FlashGraph.h:178-194 + 291-301 */
OR_EXPORT int or_is_adjacent(const or_graph* g, int32_t u, int32_t v) {
    if (iv_find(&g->added[u], v) >= 0) return 1;
    if (iv_find(&g->removed[u], v) >= 0) return 0;
    for (int64_t p = g->nindex[u]; p < g->nindex[u + 1]; p++)
        if (g->nlist[p] == v) return 1;
    return 0;
}

/* FlashGraph.h:196-203, 251-258 */
OR_EXPORT void or_insert_edge(or_graph* g, int32_t u, int32_t v) {
    int64_t i = iv_find(&g->removed[u], v);
    if (i >= 0) {
        iv_erase_at(&g->removed[u], i);
        int64_t j = iv_find(&g->removed[v], u);
        if (j >= 0) iv_erase_at(&g->removed[v], j);
    } else {
        iv_push(&g->added[u], v);
        iv_push(&g->added[v], u);
    }
    g->delta_ops++;
}

/* FlashGraph.h:205-212, 260-275 */
OR_EXPORT void or_remove_edge(or_graph* g, int32_t u, int32_t v) {
    int64_t i = iv_find(&g->added[u], v);
    if (i >= 0) {
        iv_erase_at(&g->added[u], i);
        int64_t j = iv_find(&g->added[v], u);
        if (j >= 0) iv_erase_at(&g->added[v], j);
    } else {
        if (iv_find(&g->removed[u], v) < 0) iv_push(&g->removed[u], v);
        if (iv_find(&g->removed[v], u) < 0) iv_push(&g->removed[v], u);
    }
    g->delta_ops++;
}

/* FlashGraph.h:214-233: base row minus tombstones, then additions */
static void neighbors_into(const or_graph* g, int32_t v, ivec* out) {
    out->n = 0;
    const ivec* tomb = &g->removed[v];
    int64_t b = g->nindex[v], e = g->nindex[v + 1];
    iv_reserve(out, (e - b) + g->added[v].n);
    for (int64_t p = b; p < e; p++) {
        int32_t x = g->nlist[p];
        if (tomb->n == 0 || iv_find(tomb, x) < 0) out->a[out->n++] = x;
    }
    for (int64_t i = 0; i < g->added[v].n; i++) out->a[out->n++] = g->added[v].a[i];
}

OR_EXPORT int64_t or_neighbors(const or_graph* g, int32_t v, int32_t* out, int64_t cap) {
    ivec t = {0};
    neighbors_into(g, v, &t);
    int64_t n = t.n;
    if (out) memcpy(out, t.a, (size_t)(n < cap ? n : cap) * 4);
    iv_free(&t);
    return n;
}

/* FlashGraph.h:235-242 */
OR_EXPORT int64_t or_degree(const or_graph* g, int32_t v) {
    return (g->nindex[v + 1] - g->nindex[v]) - g->removed[v].n + g->added[v].n;
}

/* F1: every parsed pair's flag is fixed against the graph as loaded
 * (misfit.cu:822, cli_setup.h:77). */
OR_EXPORT void or_flags(const or_graph* g, const int32_t* u, const int32_t* v, int64_t cnt, uint8_t* ins) {
    for (int64_t i = 0; i < cnt; i++) ins[i] = !or_is_adjacent(g, u[i], v[i]);
}

/* misfit.cu:708-714: the whole batch goes into the overlay first */
OR_EXPORT void or_apply_batch(or_graph* g, const int32_t* u, const int32_t* v, const uint8_t* ins, int64_t cnt) {
    for (int64_t i = 0; i < cnt; i++) {
        int adj = or_is_adjacent(g, u[i], v[i]);
        if (ins[i] && !adj) or_insert_edge(g, u[i], v[i]);
        else if (!ins[i] && adj) or_remove_edge(g, u[i], v[i]);
    }
}

/* ------------------------------------------------------------------ repair */
typedef struct {
    int64_t active_ins;      /* insertions with both endpoints in the MIS */
    int64_t evictions;
    int64_t admit_ins;       /* re-admissions on the insertion branch */
    int64_t active_del;      /* deletions with exactly one endpoint in the MIS */
    int64_t admit_del;       /* admissions on the deletion branch */
    int64_t completion_adds;
    int64_t seeds;
    int64_t rows;            /* seeds + candidate rows */
    int64_t row_elems;       /* packed elements (literal engine only) */
} or_stats;

enum { OR_COMPLETION = 1, OR_LITERAL = 2 };

/* Seeds in first-appearance order.  Order does not affect results. */
static int32_t* collect_seeds(const or_graph* g, const int32_t* u, const int32_t* v, int64_t cnt,
                              uint8_t* flag, int64_t* ns) {
    int32_t* s = (int32_t*)malloc((size_t)(2 * cnt + 1) * sizeof(int32_t));
    int64_t k = 0;
    for (int64_t i = 0; i < cnt; i++) {
        if (!flag[u[i]]) { flag[u[i]] = 2; s[k++] = u[i]; }
        if (!flag[v[i]]) { flag[v[i]] = 2; s[k++] = v[i]; }
    }
    (void)g;
    *ns = k;
    return s;
}

/* misfit.cu:158-202 -- row(s) = [s] ++ N1(s) ++ N2(s) in BFS discovery order.
 * Appends N1 vertices to hop1 (may contain repeats across seeds). */
static void row_2hop(const or_graph* g, int32_t s, uint8_t* visited, ivec* row, ivec* nb,
                     ivec* hop1) {
    row->n = 0;
    visited[s] = 1;
    iv_push(row, s);
    neighbors_into(g, s, nb);
    for (int64_t i = 0; i < nb->n; i++) {
        int32_t y = nb->a[i];
        if (!visited[y]) { visited[y] = 1; iv_push(row, y); }
    }
    int64_t n1_end = row->n;
    for (int64_t i = 1; i < n1_end; i++) {
        int32_t x = row->a[i];
        if (hop1) iv_push(hop1, x);
        neighbors_into(g, x, nb);
        for (int64_t j = 0; j < nb->n; j++) {
            int32_t y = nb->a[j];
            if (!visited[y]) { visited[y] = 1; iv_push(row, y); }
        }
    }
    for (int64_t i = 0; i < row->n; i++) visited[row->a[i]] = 0;
}

static int row_has_mis_except(const ivec* row, const uint8_t* mis, int32_t except) {
    for (int64_t k = 0; k < row->n; k++) {
        int32_t y = row->a[k];
        if (y != except && mis[y]) return 1;
    }
    return 0;
}

/* Rule over packed rows, misfit.cu:283-347, edges in batch order. */
static void replay_literal(const or_graph* g, uint8_t* mis, const int32_t* u, const int32_t* v,
                           const uint8_t* ins, int64_t cnt, ivec* rows, const int64_t* rowidx,
                           uint8_t* evicted, or_stats* st) {
    (void)g;
    for (int64_t e = 0; e < cnt; e++) {
        int32_t a = u[e], b = v[e];
        int ma = mis[a], mb = mis[b];
        if (ins[e]) {
            if (ma && mb) {
                st->active_ins++;
                int32_t w = a > b ? a : b;
                mis[w] = 0;
                evicted[w] = 1;
                st->evictions++;
                int64_t ri = rowidx[w];
                if (ri < 0) continue;
                const ivec* rw = &rows[ri];
                for (int64_t j = 0; j < rw->n; j++) {
                    int32_t x = rw->a[j];
                    if (mis[x]) continue;
                    int64_t xi = rowidx[x];
                    if (xi < 0) continue;
                    if (!row_has_mis_except(&rows[xi], mis, -1)) {
                        mis[x] = 1;
                        st->admit_ins++;
                        break;
                    }
                }
            }
        } else {
            if (ma && !mb) {
                st->active_del++;
                int64_t ri = rowidx[b];
                int can = 1;
                if (ri >= 0) can = !row_has_mis_except(&rows[ri], mis, a);
                if (can) { mis[b] = 1; st->admit_del++; }
            } else if (!ma && mb) {
                st->active_del++;
                int64_t ri = rowidx[a];
                int can = 1;
                if (ri >= 0) can = !row_has_mis_except(&rows[ri], mis, b);
                if (can) { mis[a] = 1; st->admit_del++; }
            }
        }
    }
}

/* ---- lazy engine: same predicates, rows never materialised ---- */
typedef struct {
    const or_graph* g;
    const uint8_t* rowflag;   /* 0 no row, 1 candidate row (adj), 2 seed row (2-hop) */
    int32_t* stamp;           /* per-vertex visit stamps */
    int32_t cur;
    ivec nb, nb2, n1;
} lazy_ctx;

static int32_t next_stamp(lazy_ctx* c) {
    if (++c->cur == INT32_MAX) {
        memset(c->stamp, 0, (size_t)c->g->n * sizeof(int32_t));
        c->cur = 1;
    }
    return c->cur;
}

/* Does row(x) contain an MIS member other than `except`? */
static int lazy_row_has_mis(lazy_ctx* c, int32_t x, const uint8_t* mis, int32_t except) {
    const or_graph* g = c->g;
    if (c->rowflag[x] == 1) {
        neighbors_into(g, x, &c->nb);
        for (int64_t i = 0; i < c->nb.n; i++) {
            int32_t y = c->nb.a[i];
            if (y != except && mis[y]) return 1;
        }
        return 0;
    }
    /* seed row: {x} ∪ N1(x) ∪ N2(x) */
    if (x != except && mis[x]) return 1;
    neighbors_into(g, x, &c->n1);
    for (int64_t i = 0; i < c->n1.n; i++) {
        int32_t y = c->n1.a[i];
        if (y != except && mis[y]) return 1;
    }
    int32_t s = next_stamp(c);
    for (int64_t i = 0; i < c->n1.n; i++) {
        int32_t y = c->n1.a[i];
        if (c->stamp[y] == s) continue;
        c->stamp[y] = s;
        neighbors_into(g, y, &c->nb);
        for (int64_t j = 0; j < c->nb.n; j++) {
            int32_t z = c->nb.a[j];
            if (z != except && mis[z]) return 1;
        }
    }
    return 0;
}

static int lazy_admissible(lazy_ctx* c, int32_t x, const uint8_t* mis) {
    if (mis[x]) return 0;
    if (!c->rowflag[x]) return 0;
    return !lazy_row_has_mis(c, x, mis, -1);
}

static void replay_lazy(lazy_ctx* c, uint8_t* mis, const int32_t* u, const int32_t* v,
                        const uint8_t* ins, int64_t cnt, uint8_t* evicted, or_stats* st) {
    const or_graph* g = c->g;
    ivec fw = {0}, nbx = {0};
    int32_t* seen = NULL;  /* per-scan memo: stamp when x was found non-admissible */
    seen = (int32_t*)calloc((size_t)g->n, sizeof(int32_t));
    int32_t sstamp = 0;
    for (int64_t e = 0; e < cnt; e++) {
        int32_t a = u[e], b = v[e];
        int ma = mis[a], mb = mis[b];
        if (ins[e]) {
            if (!(ma && mb)) continue;
            st->active_ins++;
            int32_t w = a > b ? a : b;
            mis[w] = 0;
            evicted[w] = 1;
            st->evictions++;
            /* first admissible vertex of [w] ++ adj(w) ++ concat adj(adj(w)):
             * the first occurrence order of this sequence is the BFS
             * discovery order of row(w) (misfit.cu:176-194). */
            if (++sstamp == INT32_MAX) { memset(seen, 0, (size_t)g->n * 4); sstamp = 1; }
            int32_t found = -1;
            if (lazy_admissible(c, w, mis)) found = w;
            seen[w] = sstamp;
            if (found < 0) {
                neighbors_into(g, w, &fw);
                for (int64_t i = 0; i < fw.n && found < 0; i++) {
                    int32_t x = fw.a[i];
                    if (seen[x] == sstamp) continue;
                    seen[x] = sstamp;
                    if (lazy_admissible(c, x, mis)) found = x;
                }
                for (int64_t i = 0; i < fw.n && found < 0; i++) {
                    neighbors_into(g, fw.a[i], &nbx);
                    for (int64_t j = 0; j < nbx.n; j++) {
                        int32_t y = nbx.a[j];
                        if (seen[y] == sstamp) continue;
                        seen[y] = sstamp;
                        if (lazy_admissible(c, y, mis)) { found = y; break; }
                    }
                }
            }
            if (found >= 0) { mis[found] = 1; st->admit_ins++; }
        } else {
            int32_t tgt, other;
            if (ma && !mb) { tgt = b; other = a; }
            else if (!ma && mb) { tgt = a; other = b; }
            else continue;
            st->active_del++;
            int can = 1;
            if (c->rowflag[tgt]) can = !lazy_row_has_mis(c, tgt, mis, other);
            if (can) { mis[tgt] = 1; st->admit_del++; }
        }
    }
    free(seen);
    iv_free(&fw);
    iv_free(&nbx);
}

/* Completion (this project's addition): C = S ∪ E ∪ N(E) where E is the set
 * of evicted vertices; in ascending id, x joins if it is outside the MIS and
 * has no MIS neighbour. */
static int64_t completion(const or_graph* g, uint8_t* mis, const uint8_t* rowflag,
                          const uint8_t* evicted) {
    int32_t n = g->n;
    uint8_t* inC = (uint8_t*)calloc((size_t)n, 1);
    ivec nb = {0};
    for (int32_t x = 0; x < n; x++) {
        if (rowflag[x] == 2) inC[x] = 1;
        if (evicted[x]) {
            inC[x] = 1;
            neighbors_into(g, x, &nb);
            for (int64_t i = 0; i < nb.n; i++) inC[nb.a[i]] = 1;
        }
    }
    int64_t adds = 0;
    for (int32_t x = 0; x < n; x++) {
        if (!inC[x] || mis[x]) continue;
        neighbors_into(g, x, &nb);
        int covered = 0;
        for (int64_t i = 0; i < nb.n; i++) if (mis[nb.a[i]]) { covered = 1; break; }
        if (!covered) { mis[x] = 1; adds++; }
    }
    iv_free(&nb);
    free(inC);
    return adds;
}

/* Repair one batch whose updates are ALREADY applied to the overlay. */
OR_EXPORT int or_repair(or_graph* g, uint8_t* mis, const int32_t* u, const int32_t* v,
                        const uint8_t* ins, int64_t cnt, int mode, int threads, or_stats* st_out) {
    or_stats st;
    memset(&st, 0, sizeof(st));
    int32_t n = g->n;
    uint8_t* rowflag = (uint8_t*)calloc((size_t)n, 1);
    uint8_t* evicted = (uint8_t*)calloc((size_t)n, 1);
    int64_t ns = 0;
    int32_t* seeds = collect_seeds(g, u, v, cnt, rowflag, &ns);
    st.seeds = ns;
    if (threads <= 0) threads = 1;

    if (mode & OR_LITERAL) {
        /* rows for seeds: OpenMP over seeds, as misfit.cu:532-548 */
        ivec* srows = (ivec*)calloc((size_t)(ns ? ns : 1), sizeof(ivec));
        ivec* hop1 = (ivec*)calloc((size_t)threads, sizeof(ivec));
        #pragma omp parallel num_threads(threads)
        {
            uint8_t* visited = (uint8_t*)calloc((size_t)n, 1);
            ivec nb = {0};
            int t = omp_get_thread_num();
            #pragma omp for schedule(dynamic)
            for (int64_t i = 0; i < ns; i++) row_2hop(g, seeds[i], visited, &srows[i], &nb, &hop1[t]);
            free(visited);
            iv_free(&nb);
        }
        /* candidate rows: N1(S) \ S, row = adj(x)  (misfit.cu:550-558) */
        ivec cand = {0};
        for (int t = 0; t < threads; t++) {
            for (int64_t i = 0; i < hop1[t].n; i++) {
                int32_t x = hop1[t].a[i];
                if (!rowflag[x]) { rowflag[x] = 1; iv_push(&cand, x); }
            }
            iv_free(&hop1[t]);
        }
        free(hop1);
        int64_t nrows = ns + cand.n;
        ivec* rows = (ivec*)calloc((size_t)(nrows ? nrows : 1), sizeof(ivec));
        int64_t* rowidx = (int64_t*)malloc((size_t)n * sizeof(int64_t));
        for (int32_t x = 0; x < n; x++) rowidx[x] = -1;
        for (int64_t i = 0; i < ns; i++) { rows[i] = srows[i]; rowidx[seeds[i]] = i; }
        free(srows);
        #pragma omp parallel num_threads(threads)
        {
            #pragma omp for schedule(dynamic, 64)
            for (int64_t i = 0; i < cand.n; i++) neighbors_into(g, cand.a[i], &rows[ns + i]);
        }
        for (int64_t i = 0; i < cand.n; i++) rowidx[cand.a[i]] = ns + i;
        st.rows = nrows;
        for (int64_t i = 0; i < nrows; i++) st.row_elems += rows[i].n;
        replay_literal(g, mis, u, v, ins, cnt, rows, rowidx, evicted, &st);
        for (int64_t i = 0; i < nrows; i++) iv_free(&rows[i]);
        free(rows);
        free(rowidx);
        iv_free(&cand);
    } else {
        ivec nb = {0};
        int64_t nc = 0;
        for (int64_t i = 0; i < ns; i++) {
            neighbors_into(g, seeds[i], &nb);
            for (int64_t j = 0; j < nb.n; j++) {
                int32_t x = nb.a[j];
                if (!rowflag[x]) { rowflag[x] = 1; nc++; }
            }
        }
        iv_free(&nb);
        st.rows = ns + nc;
        lazy_ctx c;
        memset(&c, 0, sizeof(c));
        c.g = g;
        c.rowflag = rowflag;
        c.stamp = (int32_t*)calloc((size_t)n, sizeof(int32_t));
        replay_lazy(&c, mis, u, v, ins, cnt, evicted, &st);
        free(c.stamp);
        iv_free(&c.nb); iv_free(&c.nb2); iv_free(&c.n1);
    }
    if (mode & OR_COMPLETION) st.completion_adds = completion(g, mis, rowflag, evicted);
    free(seeds);
    free(rowflag);
    free(evicted);
    if (st_out) *st_out = st;
    return 0;
}

/* apply + repair, the unit the binaries time per batch (misfit.cu:704-730) */
OR_EXPORT int or_process_batch(or_graph* g, uint8_t* mis, const int32_t* u, const int32_t* v,
                               const uint8_t* ins, int64_t cnt, int mode, int threads, or_stats* st) {
    or_apply_batch(g, u, v, ins, cnt);
    return or_repair(g, mis, u, v, ins, cnt, mode, threads, st);
}

/* Independence + maximality over base plus overlay. */
OR_EXPORT void or_check(const or_graph* g, const uint8_t* mis, int64_t* conflicts, int64_t* uncovered) {
    int64_t conf = 0, unc = 0;
    ivec nb = {0};
    for (int32_t x = 0; x < g->n; x++) {
        neighbors_into(g, x, &nb);
        int has = 0;
        for (int64_t i = 0; i < nb.n; i++) {
            int32_t y = nb.a[i];
            if (mis[y]) {
                has = 1;
                if (mis[x] && y != x) conf++;
            }
        }
        if (!mis[x] && !has) unc++;
    }
    iv_free(&nb);
    *conflicts = conf;   /* directed count: each conflicting edge counted twice */
    *uncovered = unc;
}

/* flash_compactor.h:20-86: row = base minus tombstones, then additions in
 * insertion order.  Returns the new entry count; fills the outputs if given. */
OR_EXPORT int64_t or_compact(const or_graph* g, int64_t* nindex_out, int32_t* nlist_out) {
    int64_t total = 0;
    for (int32_t x = 0; x < g->n; x++) total += or_degree(g, x);
    if (!nindex_out) return total;
    ivec nb = {0};
    int64_t off = 0;
    nindex_out[0] = 0;
    for (int32_t x = 0; x < g->n; x++) {
        neighbors_into(g, x, &nb);
        if (nlist_out && nb.n) memcpy(nlist_out + off, nb.a, (size_t)nb.n * 4);
        off += nb.n;
        nindex_out[x + 1] = off;
    }
    iv_free(&nb);
    return off;
}

/* tools/compute_initial_mis.py:32-43, on the current graph (base + overlay) */
OR_EXPORT void or_greedy_mis(const or_graph* g, uint8_t* mis) {
    uint8_t* blocked = (uint8_t*)calloc((size_t)g->n, 1);
    ivec nb = {0};
    for (int32_t x = 0; x < g->n; x++) {
        mis[x] = 0;
        if (blocked[x]) continue;
        mis[x] = 1;
        neighbors_into(g, x, &nb);
        for (int64_t i = 0; i < nb.n; i++) blocked[nb.a[i]] = 1;
    }
    iv_free(&nb);
    free(blocked);
}

/* ---------------------------------------------------------------- file io */
/* istream >> int: skip whitespace, optional sign, at least one digit,
 * stops at the first non-digit; fails on overflow. */
static const char* parse_int(const char* p, int32_t* out) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') p++;
    int neg = 0;
    if (*p == '+' || *p == '-') { neg = (*p == '-'); p++; }
    if (*p < '0' || *p > '9') return NULL;
    int64_t v = 0;
    while (*p >= '0' && *p <= '9') {
        v = v * 10 + (*p - '0');
        if (v > (int64_t)INT32_MAX + 1) return NULL;
        p++;
    }
    if (neg) v = -v;
    if (v > INT32_MAX || v < INT32_MIN) return NULL;
    *out = (int32_t)v;
    return p;
}

/* core/cli_setup.h:66-82 -- returns the number of kept pairs (or -1 if the
 * file cannot be opened).  u/v are 0-indexed; flags use the current graph. */
OR_EXPORT int64_t or_parse_batch_file(const char* path, const or_graph* g, int32_t* u, int32_t* v,
                                      uint8_t* ins, int64_t cap) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char* line = NULL;
    size_t lcap = 0;
    int64_t k = 0;
    while (getline(&line, &lcap, f) >= 0) {
        int32_t a, b;
        const char* p = parse_int(line, &a);
        if (!p) continue;
        p = parse_int(p, &b);
        if (!p) continue;
        a--; b--;
        if (a >= 0 && a < g->n && b >= 0 && b < g->n && a != b) {
            if (k < cap) {
                u[k] = a; v[k] = b;
                ins[k] = !or_is_adjacent(g, a, b);
            }
            k++;
        }
    }
    free(line);
    fclose(f);
    return k;
}

/* misfit.cu:778-791 -- 1-indexed ids, out-of-range ignored; card counts
 * in-range lines (duplicates included).  Returns -1 on an unparsable line. */
OR_EXPORT int or_read_mis(const char* path, int32_t n, uint8_t* mis, int64_t* card) {
    FILE* f = fopen(path, "r");
    memset(mis, 0, (size_t)n);
    *card = 0;
    if (!f) return 0;
    char* line = NULL;
    size_t lcap = 0;
    int rc = 0;
    while (getline(&line, &lcap, f) >= 0) {
        int32_t id;
        if (!parse_int(line, &id)) { rc = -1; break; }
        id -= 1;
        if (id >= 0 && id < n) { mis[id] = 1; (*card)++; }
    }
    free(line);
    fclose(f);
    return rc;
}
