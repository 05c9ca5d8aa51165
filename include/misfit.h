/*
 * misfit.h -- C ABI of the MI355X MISFIT engine (libmisfit_amd.so).
 *
 * Plain C types only: no HIP or torch types cross this boundary.  Every call
 * returns 0 on success or a negative MISFIT_E_* code; the message of the last
 * failure is available from misfit_last_error().  No call ever exits the
 * process.  Host buffers are borrowed for the duration of the call only;
 * device state belongs to the context.  A context is driven by one host
 * thread at a time (callers synchronise externally); several contexts may be
 * used from several threads.
 *
 * The reference (anonymous-artifacts/MISFIT) has no C ABI: its engine is the
 * C++ class GPUVertexClusteringGraph inside src/misfit/misfit.cu (and the
 * same class in src/misfit_async/misfit_async.cu).  Each entry point below
 * names the reference interface it replaces.
 */
#ifndef MISFIT_AMD_H
#define MISFIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MISFIT_OK 0
#define MISFIT_E_ARG (-1)        /* invalid argument */
#define MISFIT_E_HIP (-2)        /* HIP runtime failure */
#define MISFIT_E_STATE (-3)      /* call out of order (no graph, no MIS, ...) */
#define MISFIT_E_NOMEM (-4)      /* device or host allocation failed */
#define MISFIT_E_UNSUPPORTED (-5)/* input the engine does not handle */
#define MISFIT_E_IO (-6)         /* file input/output failure */
#define MISFIT_E_CHECK (-7)      /* an internal consistency check failed */

typedef struct misfit_ctx misfit_ctx;
typedef struct misfit_staged misfit_staged;

/* One edge update.  Vertices are 0-indexed.  is_insertion follows the
 * reference's flag: computed once against the graph as loaded
 * (src/misfit/misfit.cu:822, core/cli_setup.h:77) and then applied as
 * "insert if absent" / "delete if present" (src/misfit/misfit.cu:708-714). */
typedef struct {
    int32_t u;
    int32_t v;
    uint8_t is_insertion;
    uint8_t pad[3];
} misfit_update;

/* Per-batch statistics (counts from the device, times from the device clock). */
typedef struct {
    int64_t updates;          /* updates in the batch */
    int64_t seeds;            /* distinct endpoints */
    int64_t active_ins;       /* insertions with both ends in the MIS */
    int64_t evictions;
    int64_t admit_ins;        /* re-admissions on the insertion branch */
    int64_t active_del;       /* deletions with exactly one end in the MIS */
    int64_t admit_del;        /* admissions on the deletion branch */
    int64_t completion_adds;  /* vertices added by the completion pass */
    int64_t rounds;           /* speculative rounds of the exact replay */
    int64_t evaluations;      /* edge evaluations over all rounds */
    int64_t maybe_uncovered;  /* vertices the batch may uncover (PU) */
    int64_t admissible;       /* ... of which have a row */
    int64_t delta_ops;        /* effective overlay operations */
    int64_t overlay_entries;  /* pool slots in use after the batch */
    double device_ms;         /* batch time on the device (device wall clock, first kernel to end of tail) */
    double prologue_ms;       /* ... of which the overlay update, may-uncover set and admissible index */
    int64_t window;           /* edges speculated per round (the rounds' window) */
    double device_start_ms;   /* the device wall clock at the batch's start and end, in ms (comparable */
    double device_end_ms;     /* across the batches of every context on the device) */
    double rounds_ms;         /* of device_ms: the speculative rounds, grid-wide and in the tail */
    double counters_ms;       /* ... the rule counters and the completion's candidates (the rest after the
                                 rounds is the completion and the publication) */
} misfit_batch_stats;

/* Execution modes of the repair (results are identical). */
#define MISFIT_MODE_PLAIN 0   /* one batch at a time: the host drives every round */
#define MISFIT_MODE_QUEUED 1  /* whole batches enqueued, device-driven rounds (default) */

/* Replaces the GPUVertexClusteringGraph constructor's device selection and
 * stream creation (src/misfit/misfit.cu:389-421, core/cli_setup.h:43-47). */
int misfit_ctx_create(int device, misfit_ctx** out);

/* Replaces ~GPUVertexClusteringGraph (src/misfit/misfit.cu:423-439). */
int misfit_ctx_destroy(misfit_ctx* ctx);

/* Message of the last failed call on ctx (never NULL). */
const char* misfit_last_error(const misfit_ctx* ctx);

/* Streams the base CSR into HBM once.  nindex has n+1 entries, int64 when
 * nindex_is_64 else int32; nlist has m int32 entries.  Replaces the mmap'd
 * FlashGraph reads of the reference (core/FlashGraph.h:21-170), which the
 * engine no longer needs after this call.  Resets the overlay and the MIS. */
int misfit_load_csr(misfit_ctx* ctx, int32_t n, int64_t m, const void* nindex, int nindex_is_64,
                    const int32_t* nlist);

/* Sets the MIS (one byte per vertex, nonzero = member).  Replaces
 * GPUVertexClusteringGraph::setMISFromHost (src/misfit/misfit.cu:448-453). */
int misfit_set_mis(misfit_ctx* ctx, const uint8_t* mis);

/* Reads the MIS back.  Replaces getMISToHost (src/misfit/misfit.cu:455-462). */
int misfit_get_mis(misfit_ctx* ctx, uint8_t* mis);

/* completion_on = 1 (default): after the reference rule, add the greedy
 * ascending-id completion that keeps the MIS maximal; 0: the literal rule
 * alone.  The reference has only the literal rule (src/misfit/misfit.cu:
 * 277-349). */
int misfit_set_rule(misfit_ctx* ctx, int completion_on);

/* MISFIT_MODE_PLAIN or MISFIT_MODE_QUEUED. */
int misfit_set_mode(misfit_ctx* ctx, int mode);

/* Insertion flags against the current graph: ins[i] = !adjacent(u[i], v[i]).
 * Replaces the HybridGraph::isAdjacent call of batch parsing
 * (core/cli_setup.h:77, src/misfit/misfit.cu:822). */
int misfit_flags(misfit_ctx* ctx, const int32_t* u, const int32_t* v, int64_t count, uint8_t* ins);

/* misfit_flags without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (serving loops that parse
 * batches as they arrive, harnesses that forbid hipMalloc).  The pairs are
 * staged in a static device array of MISFIT_NOALLOC_MAX_PAIRS pairs (9 MiB,
 * held by every GPU the library runs on until the process ends); larger
 * counts return MISFIT_E_ARG.  Calls on one GPU share that array: the
 * library serialises them within a process (the array keeps the last
 * call's contents), and two processes each have their own.  Results equal
 * misfit_flags'. */
#define MISFIT_NOALLOC_MAX_PAIRS (1 << 20)
int misfit_flags_noalloc(misfit_ctx* ctx, const int32_t* u, const int32_t* v, int64_t count, uint8_t* ins);

/* Queues one batch: overlay update, exact repair, completion.  Returns at
 * once in queued mode (the host buffer is copied before the call returns).
 * The speculative rounds are queued from the recent rounds per update of
 * batches at least a quarter as large; without such history (a graph's
 * first batches, or a much larger batch) the batch may need rounds queued
 * later from the host: misfit_poll / misfit_wait do that, and the next
 * submit on the context (or any call that reads the graph) first waits for
 * such a batch to finish.  Replaces the async engine's launchBatch
 * (src/misfit_async/misfit_async.cu:484-675) together with the overlay
 * update of launchBatchGPU (:718-740). */
int misfit_submit_batch(misfit_ctx* ctx, const misfit_update* ups, int64_t count, uint64_t* ticket);

/* The host-to-device copy of a batch, taken out of the submission: the
 * reference stages every batch from pinned host memory to the device before
 * its kernels run (src/misfit/misfit.cu:129-149,612-619); here
 * misfit_stage_batch checks the updates as misfit_submit_batch does and
 * uploads them to device memory owned by the context, and
 * misfit_submit_staged queues such a batch with no host copy (the same
 * results as misfit_submit_batch with the same updates).  A staged batch may
 * be submitted any number of times, on its own context only.
 * misfit_release_staged waits for the context's stream and frees it; what
 * is not released is freed by misfit_ctx_destroy. */
int misfit_stage_batch(misfit_ctx* ctx, const misfit_update* ups, int64_t count, misfit_staged** out);
int misfit_submit_staged(misfit_ctx* ctx, const misfit_staged* batch, uint64_t* ticket);
int misfit_release_staged(misfit_ctx* ctx, misfit_staged* batch);

/* Waits for a submitted batch and returns its statistics (st may be NULL).
 * Replaces syncAndCollect (src/misfit_async/misfit_async.cu:677-701). */
int misfit_wait(misfit_ctx* ctx, uint64_t ticket, misfit_batch_stats* st);

/* Non-blocking: 1 when the submitted batch has finished on the device (its
 * results can be collected with misfit_wait without blocking), 0 while it is
 * still queued or running (a batch that needs more rounds than were queued
 * gets them here).  Replaces the reference's blocking event waits in
 * syncAndCollect (src/misfit_async/misfit_async.cu:677-701) for callers that
 * multiplex many graphs on one thread. */
int misfit_poll(misfit_ctx* ctx, uint64_t ticket);

/* submit + wait.  Replaces processBatchGPUVertexClustering /
 * processBatchWithVertexClustering (src/misfit/misfit.cu:468-730). */
int misfit_process_batch(misfit_ctx* ctx, const misfit_update* ups, int64_t count, misfit_batch_stats* st);

/* Waits for all queued work.  Replaces synchronizeChannels
 * (src/misfit/misfit.cu:464-466). */
int misfit_sync(misfit_ctx* ctx);

/* Device time of a region of the context's stream, by a pair of HIP events:
 * misfit_timer_begin records the first where the stream stands now;
 * misfit_timer_end records the second, waits for it and returns the elapsed
 * milliseconds.  A measurement aid in the place of the reference's per-stage
 * event pairs (src/misfit/misfit.cu:612-650): per-batch events on the stream
 * cost the device several microseconds each, so batch times come from the
 * device clock (misfit_batch_stats.device_ms) and whole runs from these. */
int misfit_timer_begin(misfit_ctx* ctx);
int misfit_timer_end(misfit_ctx* ctx, double* ms);

/* Full check over base plus overlay: conflicts = edges with both ends in the
 * MIS, uncovered = non-members without an MIS neighbour, count_errors =
 * vertices whose maintained MIS-neighbour count disagrees with a recount
 * (always 0 unless the engine is broken).  No reference counterpart; the
 * reference leaves checking to an external baseline
 * (src/misfit/misfit.cu:890-891). */
int misfit_check(misfit_ctx* ctx, int64_t* conflicts, int64_t* uncovered, int64_t* count_errors);

/* Merges the overlay into a new base on the device (base order minus
 * deletions, then insertions in insertion order) and clears the overlay.
 * Replaces compact_flash_graph + remap + delta.clear
 * (core/flash_compactor.h:20-86, src/misfit/misfit.cu:916-923); the file
 * write is misfit_write_egr below. */
int misfit_compact(misfit_ctx* ctx, int64_t* new_m);

/* Current graph size: n, base entries, live entries (base + overlay). */
int misfit_graph_size(misfit_ctx* ctx, int32_t* n, int64_t* base_m, int64_t* live_m);

/* Copies the base CSR (call misfit_compact first to include the overlay) to
 * host buffers of n+1 int64 and base_m int32. */
int misfit_export_csr(misfit_ctx* ctx, int64_t* nindex, int32_t* nlist);

/* Neighbours of v in the current graph, in the reference's order
 * (core/FlashGraph.h:214-233).  Returns the count through *count; writes at
 * most cap entries. */
int misfit_neighbors(misfit_ctx* ctx, int32_t v, int32_t* out, int64_t cap, int64_t* count);

/* Algorithmic bytes of the most recent batch on the current graph:
 * 16*U + sum over V1 of (I + 4*deg) + |V2|, V1 = S ∪ N(S), V2 = V1 ∪ N(V1),
 * I = index_bytes (8 for an int32 .egr, 16 for int64).  Measurement aid. */
int misfit_batch_bytes(misfit_ctx* ctx, int index_bytes, int64_t* bytes, int64_t* v1, int64_t* v2);

/* Device name and memory (bytes).  Replaces the cudaGetDeviceProperties /
 * cudaMemGetInfo calls of the reference's setup and peak-memory report
 * (src/misfit/misfit.cu:46-67,758-761). */
int misfit_device_info(misfit_ctx* ctx, char* name, int namelen, int64_t* total_bytes, int64_t* free_bytes);

/* ---- serving: many graphs on one GPU, batches taken as they arrive ----
 * Generalises the async engine's pipelining (src/misfit_async/
 * misfit_async.cu:484-701,827-876): graphs are admitted by free device
 * memory, each graph's requests run in submission order with up to
 * max_inflight batches queued on the device, batches of different graphs
 * overlap on the GPU (fed by a pool of dispatcher threads: a launch costs
 * the host about 4 us, and separate threads on separate streams scale), a
 * graph is compacted when its overlay passes
 * compact_fraction of its base entries, and every request reports its
 * queueing time and latency.  Thread-safe: any thread may submit or wait. */
typedef struct misfit_server misfit_server;

typedef struct {
    int64_t hbm_reserve;      /* device bytes kept free: graphs and batches that would cut into it are refused */
    double compact_fraction;  /* compact a graph when its overlay exceeds this share of its base (0: never) */
    int max_inflight;         /* batches per graph queued on the device at once (1..3); with 1, every
                                 batch is queued with the rounds it is expected to need and gets more
                                 from the host if it needs them, so no rounds run empty (many graphs) */
    int workers;              /* dispatcher threads, each owning a share of the graphs (0: 8) */
    int persistent;           /* with max_inflight 1: 0 every batch on its graph's stream; 1 small batches,
                                 contended graphs' batches and first batches run whole, each in one
                                 workgroup of one long-running serving kernel shared by all graphs (its
                                 workgroups take batches from a queue as they free up), the rest on their
                                 streams; 2 every batch in the serving kernel */
} misfit_server_opts;

typedef struct {
    int32_t graph;
    int64_t updates;
    int64_t compactions_before;  /* compactions of the graph that ran before this batch started */
    double queue_ms;             /* submission -> queued on the device */
    double latency_ms;           /* submission -> result */
    double device_ms;            /* the batch on the device */
    misfit_batch_stats batch;
} misfit_request_stats;

int misfit_server_create(int device, const misfit_server_opts* opts, misfit_server** out);
int misfit_server_destroy(misfit_server* srv);
const char* misfit_server_last_error(const misfit_server* srv);
/* MISFIT_E_NOMEM (nothing allocated) when the graph does not fit: its
 * resident bytes (base, sorted row copy, overlay, per-vertex state) plus the
 * headroom of the largest compaction a resident graph may run (compactions
 * run one at a time) must fit in free device memory minus hbm_reserve.  With
 * MISFIT_SERVE_COMPACT_KEEP=1 (read when the server is created) every graph
 * keeps its compactions' buffers and is charged its own headroom instead. */
int misfit_server_add_graph(misfit_server* srv, int32_t n, int64_t m, const void* nindex, int nindex_is_64,
                            const int32_t* nlist, const uint8_t* mis, int32_t* graph);
int misfit_server_remove_graph(misfit_server* srv, int32_t graph);
/* Changes the device bytes kept free (hbm_reserve) for every later
 * admission of graphs, batches and compactions, e.g. to make room for
 * another tenant of the GPU.  No reference counterpart: the reference
 * serves one graph and sizes nothing by free memory. */
int misfit_server_set_reserve(misfit_server* srv, int64_t hbm_reserve);
/* Returns at once; the updates are copied. */
int misfit_server_submit(misfit_server* srv, int32_t graph, const misfit_update* ups, int64_t count,
                         uint64_t* request);
/* Blocks until the request has finished; its result code (MISFIT_E_NOMEM
 * when the batch was not admitted; the graph is then unchanged).  A graph
 * whose compaction could not run (its headroom does not fit in free device
 * memory minus the reserve, or the compaction failed) stays as it was, and
 * the next of its requests to start fails with that error without being
 * applied; the compaction is tried again once the overlay has grown by
 * another compact_fraction of the base. */
int misfit_server_wait(misfit_server* srv, uint64_t request, misfit_request_stats* st);
/* After waiting for a graph's requests: its MIS, and the full check. */
int misfit_server_get_mis(misfit_server* srv, int32_t graph, uint8_t* mis);
int misfit_server_check(misfit_server* srv, int32_t graph, int64_t* conflicts, int64_t* uncovered,
                        int64_t* count_errors);

/* ---- host I/O (same library, no device needed) ---- */

/* .egr reader (core/FlashGraph.h:108-169 format; both nindex widths).  The
 * file is memory-mapped; the returned pointers stay valid until
 * misfit_egr_close. */
typedef struct {
    int32_t n;
    int64_t m;
    int wide;               /* 1: int64 nindex (-1 sentinel form) */
    const void* nindex;     /* n+1 entries, int64 if wide else int32 */
    const int32_t* nlist;   /* m entries */
    void* priv;
} misfit_egr;

int misfit_egr_open(const char* path, misfit_egr* out, char* err, int errlen);
int misfit_egr_close(misfit_egr* g);

/* Writes the wide form atomically: <path>.tmp.<pid>, fflush, fsync, rename,
 * directory fsync (core/flash_compactor.h:40-104). */
int misfit_write_egr(const char* path, int32_t n, int64_t m, const int64_t* nindex, const int32_t* nlist,
                     char* err, int errlen);

/* .mtx -> .egr conversion.  Replaces tools/efficient_converter.py:9-113 with
 * the same result: the first non-comment line is the dimension line (n = the
 * larger of its first two fields); every later line "u v" (1-indexed) with
 * u <= n appends v-1 to row u-1 in file order; lines with u > n are dropped;
 * nothing is symmetrised, sorted or deduplicated; the wide form is written.
 * Returns n and the number of entries written. */
int misfit_mtx_to_egr(const char* mtx_path, const char* egr_path, int32_t* n, int64_t* m, char* err, int errlen);

/* .egr validator (no reference counterpart: the reference trusts its input,
 * core/FlashGraph.h:108-169).  Counts offsets that are not a valid prefix
 * sum, entries outside [0, n), self-loops, repeated entries in a row, rows
 * not strictly ascending and, on ascending in-range rows, entries (u, v)
 * without (v, u) (-1 when that is not checked). */
typedef struct {
    int32_t n;
    int64_t m;
    int wide;
    int64_t bad_offsets;
    int64_t out_of_range;
    int64_t self_loops;
    int64_t duplicates;
    int64_t unsorted_rows;
    int64_t asymmetric;
} misfit_egr_report;

int misfit_egr_check(const char* path, misfit_egr_report* report, char* err, int errlen);

/* ---- synthetic workloads (seeded, deterministic) ---- */

/* Erdős–Rényi G(n, M): M distinct undirected pairs, symmetric CSR with rows
 * ascending.  Returns entries via *m; call with nindex = NULL to size. */
int misfit_gen_er(int32_t n, int64_t M, uint64_t seed, int64_t* nindex, int32_t* nlist, int64_t* m);

/* R-MAT (Chakrabarti, Zhan, Faloutsos 2004): 2^scale vertices,
 * edgefactor*2^scale generated edges, quadrant probabilities a,b,c (d = 1-a-b-c),
 * self-loops dropped, symmetrised, deduplicated, rows ascending. */
int misfit_gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                    int64_t* nindex, int32_t* nlist, int64_t* m);

/* 2-D random geometric graph: n points uniform in the unit square, an edge
 * when the distance is below r = sqrt(avg_degree / (pi n)). */
int misfit_gen_rgg(int32_t n, double avg_degree, uint64_t seed, int64_t* nindex, int32_t* nlist, int64_t* m);

/* Greedy MIS in ascending id on a CSR (tools/compute_initial_mis.py:32-43). */
int misfit_greedy_mis(int32_t n, const int64_t* nindex, const int32_t* nlist, uint8_t* mis);

/* Update streams (u, v 0-indexed), written to u/v arrays of count entries.
 *   kind 0: uniform vertex pairs u != v
 *   kind 1: half deletions of uniform existing edges, half insertions of
 *           uniform non-edges, distinct across the whole stream, shuffled
 *   kind 2: both endpoints drawn proportionally to (degree + 1)
 *   kind 3: endpoints drawn from the lowest-degree quarter of the vertices
 * and the reference's batch samplers (distinct existing edges as (lo, hi),
 * shuffled; count must not exceed the entries the graph can supply):
 *   kind 4: a uniform CSR entry (tools/generate_test_batches.py:55-78)
 *   kind 5: a vertex drawn proportionally to 1/degree, then a uniform
 *           neighbour (tools/generate_low_degree_batches.py:56-100)
 *   kind 6: a vertex drawn proportionally to its degree, then a uniform
 *           neighbour (tools/generate_preferential_batches.py:56-98) */
int misfit_gen_updates(int kind, int32_t n, const int64_t* nindex, const int32_t* nlist, int64_t count,
                       uint64_t seed, int32_t* u, int32_t* v);

#ifdef __cplusplus
}
#endif
#endif
