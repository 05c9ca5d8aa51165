// capi.hip -- the C ABI (include/misfit.h): context, device memory,
// capacity management and the two ways of driving a batch.
//
//   plain   per-phase kernels; the host reads the round outcome after every
//           speculative round and every completion iteration (one batch at
//           a time, the way the reference's sync engine drives its kernels).
//   queued  the round state lives on the device, so a batch is enqueued in
//           full without host round trips: the prologue, a few grid-wide
//           rounds (no-ops once the batch is done) and one single-workgroup
//           kernel that finishes the remaining rounds and the completion.
//           submit returns at once and batches queue back to back while the
//           host stages the next ones (the reference's misfit_async
//           pipelining, generalised).  Kernel boundaries are the grid-wide
//           barrier: on a multi-die MI355X a stream launch costs ~2.4 us,
//           an in-kernel grid barrier 14-80 us.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "graph.h"
#include "../../include/misfit.h"

// A named range for rocprofv3 --marker-trace (no cost beyond a call when no
// tool is attached).  Host-side: in plain mode, where the host waits for
// each round, a range spans the device work it launched; in queued mode it
// spans the enqueue.
namespace {
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

// Consecutive roctx ranges of one call's host phases: to() ends the phase
// before it and starts the next; the last one ends with the object.
struct Phases {
    bool open = false;
    void to(const char* name) {
        if (open) roctxRangePop();
        roctxRangePushA(name);
        open = true;
    }
    ~Phases() {
        if (open) roctxRangePop();
    }
};
}  // namespace

#define MF_NSLOT 4

struct Slot {
    int32_t* h_u = nullptr;
    int32_t* h_v = nullptr;
    uint8_t* h_ins = nullptr;
    int32_t* d_u = nullptr;
    int32_t* d_v = nullptr;
    uint8_t* d_ins = nullptr;
    u64* h_ctr = nullptr;
    const int32_t* g_u = nullptr;  // the staging buffers as the device addresses them
    const int32_t* g_v = nullptr;
    const uint8_t* g_ins = nullptr;
    hipEvent_t done = nullptr;
    int64_t count = 0;
    uint64_t ticket = 0;
    bool busy = false;
    BP P{};                // the batch's parameters and grids, for rounds queued after its tail halted
    LaunchCfg L{0, nullptr};
    int par_next = 0;      // parity of the batch's next round and tail
    bool may_halt = false; // queued without a reliable round estimate; not yet settled
    bool external = false; // handed to the serving kernel: done when h_ctr[MF_DONE_WORD] is set
};

// A batch handed to the serving kernel has finished once the kernel wrote
// its item number (plus one) after the counters.
static bool external_done(const Slot& s) { return __atomic_load_n(&s.h_ctr[MF_DONE_WORD], __ATOMIC_ACQUIRE) != 0; }

struct misfit_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    GraphState S{};
    bool has_graph = false;
    int completion = 1;
    bool maximal = false;  // the MIS has no uncovered non-member (checked at set_mis, kept by the completion)
    bool force_pu_scan = false;  // MISFIT_PU_SCAN=full: always scan every vertex for PU
    // MISFIT_PU_DENSE_FRACTION=f: the may-uncover scan switches from the
    // tagged vertices only (sparse) to every vertex (dense) when the batch
    // may tag more than f*n vertices (its updates times the mean degree);
    // below 0, never
    double pu_dense_fraction = -1.0;
    int mode = MISFIT_MODE_QUEUED;
    int grid_rounds = 0;       // MISFIT_GRID_ROUNDS: fixed number of queued grid rounds
    bool grid_rounds_fixed = false;
    double hist_rpu[16] = {0};    // (rounds - 1) per update of recent batches of >= 512 updates
    int64_t hist_count[16] = {0}; // ... and their sizes
    int hist_pos = 0;
    bool always_halt = false;     // every batch may halt (nothing is ever queued behind one)
    double window_factor = 16.0;  // MISFIT_WINDOW_FACTOR: window = this x expected edges per round
    int round_margin = -1;        // MISFIT_ROUND_MARGIN: rounds queued beyond the planned rate (-1: by the plan)
    int fin_window = 0;           // MISFIT_TAIL_WINDOW: edges per round of the one-workgroup tail (0: 1024)
    int serve_window = 0;         // MISFIT_SERVE_WINDOW: edges per round of batches in the serving kernel (0: adaptive)
    double serve_window_factor = 2.0;  // MISFIT_SERVE_WINDOW_FACTOR: ... adaptive: this x the last edges per round
    double last_epr = 0.0;        // edges committed per round by the last batch (0: none yet)
    bool fail_growth = false;     // MISFIT_TEST_FAIL_GROWTH=1: every capacity growth fails (fault injection)
    double clock_khz = 0.0;       // device wall clock, for per-batch device times
    hipEvent_t timer0 = nullptr, timer1 = nullptr;  // misfit_timer_begin/end
    int round_per_cu = 0;      // MISFIT_ROUND_WG_PER_CU: round grid (0: by occupancy)
    // MISFIT_ROUND_THREADS: workgroup size of the round kernels (128, 256 or
    // 512); 0 (the default) chooses per batch: 512 when a round is expected
    // to commit few enough edges for the 512-thread grid (fewer, larger
    // workgroups were 2.7% faster on R-MAT scale 22's 10k-update batches),
    // else 256, whose grid holds more waves (50k-update batches on the RGG,
    // about 2,900 edges a round, ran 25 rounds instead of 17 at 512)
    int round_threads = 0;
    bool debug_rounds = false;
    uint32_t epoch = 0;
    uint32_t bytes_tag = 0;
    int64_t last_U = 0;
    // capacities
    int64_t U_cap = 0;
    u64 bt_size = 0;
    int64_t fa_cap = 0;
    int64_t ph_cap = 0;
    u64* d_ktask = nullptr;    // row-chunk tasks, sized for the graph plus the pool
    int32_t* d_sval = nullptr; // sorted copy of the base rows when they are not ascending
    int32_t* d_spos = nullptr;
    // compaction scratch, kept between compactions (grown as needed, freed
    // with the graph): the plan scan's block sums, per-vertex offsets and
    // row lists, and the tail lists of the merge
    std::vector<void*> cs_allocs;
    int64_t cs_n = 0, cs_t = 0;
    int64_t *cs_part = nullptr, *cs_doff = nullptr, *cs_toff = nullptr;
    int32_t *cs_trows = nullptr, *cs_big = nullptr;
    u64* cs_cnt = nullptr;  // [0] rows with a tail, [1] long lists
    int32_t *cs_tval = nullptr, *cs_tidx = nullptr, *cs_tsv = nullptr, *cs_tsi = nullptr;
    bool compact_selfcheck = false;  // MISFIT_COMPACT_SELFCHECK=1: full checks of every compaction's output
    void* cs_tmp = nullptr;
    size_t cs_tmp_bytes = 0;
    // MISFIT_COMPACT_KEEP (default 1): the buffers a compaction releases (the
    // old base rows and offsets, the old sorted copy, the deletion marks) are
    // kept as the next compaction's outputs instead of being freed, since
    // freeing and allocating them cost about 0.3 ms per GB; fresh compaction
    // buffers get 1/32 more room, so the next, slightly larger output still
    // fits.  The graph then holds about twice its base arrays.  Only the
    // latest compaction's buffers are kept, and all of them go at reload,
    // destroy or a failed allocation.  The serving layer turns it off: its
    // graphs share one compaction headroom.
    bool keep_compact = true;
    struct Spare {
        void* p;
        size_t cap;
        uint64_t gen;  // the compaction that released it
    };
    std::vector<Spare> spares;
    uint64_t compact_gen = 0;
    size_t cap_nindex = 0, cap_nlist = 0, cap_sval = 0, cap_spos = 0;  // bytes held by the graph's arrays
    int64_t ktask_cap = 0;
    int64_t ins_bound = 0;    // inserted pairs since the overlay was last cleared (upper bound)
    int64_t pool_bound = 0;   // pool slots possibly used (upper bound)
    // owned device memory
    std::vector<void*> graph_allocs;   // freed on reload
    std::vector<void*> batch_allocs;   // freed on batch-capacity growth
    int64_t* d_nindex = nullptr;
    int32_t* d_nlist = nullptr;
    BP P{};
    u64* d_ctr = nullptr;
    uint32_t* d_round = nullptr;
    u64* d_acc = nullptr;
    int* d_flags = nullptr;
    u64* h_small = nullptr;    // pinned scratch for small readbacks
    Slot slots[MF_NSLOT];
    uint64_t next_ticket = 1;
    std::vector<misfit_staged*> staged;  // batches uploaded ahead of submission, freed on destroy
};

// A batch uploaded to device memory ahead of its submission.
struct misfit_staged {
    int32_t* u = nullptr;
    int32_t* v = nullptr;
    uint8_t* ins = nullptr;
    int64_t count = 0;
    int64_t nins = 0;
    int32_t n = 0;  // the vertex count the endpoints were checked against
};

// ---------------------------------------------------------------- frees
namespace {
std::mutex g_free_mu;
int g_serving = 0;                 // serving kernels that may be running
std::vector<void*> g_dev_deferred;  // frees waiting for them to end
std::vector<void*> g_host_deferred;
}  // namespace

void mf_free_device(void* p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_free_mu);
        if (g_serving > 0) {
            g_dev_deferred.push_back(p);
            return;
        }
    }
    hipFree(p);
}

void mf_free_host(void* p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_free_mu);
        if (g_serving > 0) {
            g_host_deferred.push_back(p);
            return;
        }
    }
    hipHostFree(p);
}

void mf_serving_begin() {
    std::lock_guard<std::mutex> lk(g_free_mu);
    g_serving++;
}

void mf_serving_end() {
    std::vector<void*> dev, host;
    {
        std::lock_guard<std::mutex> lk(g_free_mu);
        if (--g_serving > 0) return;
        g_serving = 0;
        dev.swap(g_dev_deferred);
        host.swap(g_host_deferred);
    }
    for (void* p : dev) hipFree(p);
    for (void* p : host) hipHostFree(p);
}

// ---------------------------------------------------------------- errors
static int fail(misfit_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                         \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail((c), MISFIT_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e));    \
    } while (0)

// A copy on the context's stream, then a wait for that stream: every copy is
// ordered with the context's kernels on both sides.  (The null stream of a
// plain hipMemcpy does not order against the context's non-blocking stream,
// and a device-to-device or pageable host-to-device hipMemcpy may return
// before its copy has landed.)
static hipError_t copy_sync(misfit_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

static int build_row_index(misfit_ctx* c);
static void free_compact_scratch(misfit_ctx* c);
static int settle(misfit_ctx* c);
static hipError_t resolve_scratch(int device);

template <typename T>
static hipError_t dalloc(std::vector<void*>& owner, T** p, int64_t count) {
    size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipSuccess) owner.push_back((void*)*p);
    return e;
}

static void free_list(std::vector<void*>& v) {
    for (void* p : v) mf_free_device(p);
    v.clear();
}

static u64 pow2_at_least(u64 x) {
    u64 p = 1;
    while (p < x) p <<= 1;
    return p;
}

extern "C" const char* misfit_last_error(const misfit_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

// ---------------------------------------------------------------- context
extern "C" int misfit_ctx_create(int device, misfit_ctx** out) {
    if (!out) return MISFIT_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MISFIT_E_HIP;
    if (device < 0 || device >= ndev) return MISFIT_E_ARG;
    misfit_ctx* c = new misfit_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MISFIT_E_HIP;
    }
    if (resolve_scratch(device) != hipSuccess) {  // misfit_flags_noalloc's staging on this device
        hipStreamDestroy(c->stream);
        delete c;
        return MISFIT_E_HIP;
    }
    const char* dbg = getenv("MISFIT_DEBUG_ROUNDS");
    c->debug_rounds = dbg && dbg[0] == '1';
    const char* rw = getenv("MISFIT_ROUND_WG_PER_CU");
    if (rw) c->round_per_cu = std::max(0, std::min(8, atoi(rw)));
    const char* rt = getenv("MISFIT_ROUND_THREADS");
    if (rt && (atoi(rt) == 128 || atoi(rt) == 256 || atoi(rt) == 512)) c->round_threads = atoi(rt);
    const char* wf = getenv("MISFIT_WINDOW_FACTOR");
    if (wf && atof(wf) >= 1.0) c->window_factor = atof(wf);
    const char* rm = getenv("MISFIT_ROUND_MARGIN");
    if (rm) c->round_margin = std::max(0, std::min(16, atoi(rm)));
    const char* ps = getenv("MISFIT_PU_SCAN");
    c->force_pu_scan = ps && strcmp(ps, "full") == 0;
    const char* pd = getenv("MISFIT_PU_DENSE_FRACTION");
    if (pd) c->pu_dense_fraction = atof(pd);
    const char* sc = getenv("MISFIT_COMPACT_SELFCHECK");
    c->compact_selfcheck = sc && sc[0] == '1';
    const char* ck = getenv("MISFIT_COMPACT_KEEP");
    if (ck) c->keep_compact = ck[0] != '0';
    const char* fg = getenv("MISFIT_TEST_FAIL_GROWTH");
    c->fail_growth = fg && fg[0] == '1';
    const char* swf = getenv("MISFIT_SERVE_WINDOW_FACTOR");
    if (swf && atof(swf) > 0.0) c->serve_window_factor = atof(swf);
    const char* sw = getenv("MISFIT_SERVE_WINDOW");
    if (sw) c->serve_window = std::max(0, std::min(1 << 20, atoi(sw)));
    const char* tw = getenv("MISFIT_TAIL_WINDOW");
    if (tw) c->fin_window = std::max(0, std::min(1 << 20, atoi(tw)));
    const char* gr = getenv("MISFIT_GRID_ROUNDS");
    if (gr) {
        c->grid_rounds = std::max(0, atoi(gr));
        c->grid_rounds_fixed = true;
    }
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    c->S.blocks = std::max(64, cus * 8);
    c->S.stream = c->stream;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->clock_khz = (double)khz;
    hipEventCreate(&c->timer0);
    hipEventCreate(&c->timer1);
    if (hipHostMalloc((void**)&c->h_small, C_NCOUNTERS * MF_CS * sizeof(u64), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return MISFIT_E_NOMEM;
    }
    for (auto& s : c->slots) {
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        // written by the batch's last kernel, read after its event
        hipHostMalloc((void**)&s.h_ctr, MF_HOUT_WORDS * sizeof(u64), hipHostMallocCoherent);
    }
    *out = c;
    return MISFIT_OK;
}

// Staging buffers only: a slot's ticket, counters and event stay, so a
// batch submitted before the buffers were regrown can still be waited on
// (the regrowth synchronises the stream first).
static void free_slots(misfit_ctx* c) {
    for (auto& s : c->slots) {
        if (s.h_u) mf_free_host(s.h_u);
        if (s.h_v) mf_free_host(s.h_v);
        if (s.h_ins) mf_free_host(s.h_ins);
        s.h_u = nullptr; s.h_v = nullptr; s.h_ins = nullptr;
        s.g_u = nullptr; s.g_v = nullptr; s.g_ins = nullptr;
        s.d_u = nullptr; s.d_v = nullptr; s.d_ins = nullptr;
    }
}

extern "C" int misfit_ctx_destroy(misfit_ctx* c) {
    if (!c) return MISFIT_E_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    free_slots(c);
    for (auto& s : c->slots) {
        if (s.h_ctr) mf_free_host(s.h_ctr);
        hipEventDestroy(s.done);
    }
    hipEventDestroy(c->timer0);
    hipEventDestroy(c->timer1);
    free_list(c->batch_allocs);
    free_list(c->graph_allocs);
    free_compact_scratch(c);
    for (misfit_staged* b : c->staged) {
        mf_free_device(b->u);
        delete b;
    }
    if (c->d_ktask) mf_free_device(c->d_ktask);
    if (c->d_sval) mf_free_device(c->d_sval);
    if (c->d_spos) mf_free_device(c->d_spos);
    if (c->h_small) mf_free_host(c->h_small);
    hipStreamDestroy(c->stream);
    delete c;
    return MISFIT_OK;
}

extern "C" int misfit_set_rule(misfit_ctx* c, int completion_on) {
    if (!c) return MISFIT_E_ARG;
    c->completion = completion_on ? 1 : 0;
    return MISFIT_OK;
}

extern "C" int misfit_set_mode(misfit_ctx* c, int mode) {
    if (!c) return MISFIT_E_ARG;
    if (mode != MISFIT_MODE_PLAIN && mode != MISFIT_MODE_QUEUED) return fail(c, MISFIT_E_ARG, "unknown mode");
    c->mode = mode;
    return MISFIT_OK;
}

extern "C" int misfit_sync(misfit_ctx* c) {
    if (!c) return MISFIT_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MISFIT_OK;
}

extern "C" int misfit_timer_begin(misfit_ctx* c) {
    if (!c) return MISFIT_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->timer0, c->stream));
    return MISFIT_OK;
}

extern "C" int misfit_timer_end(misfit_ctx* c, double* ms) {
    if (!c || !ms) return MISFIT_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipEventRecord(c->timer1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->timer1));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, c->timer0, c->timer1));
    *ms = (double)f;
    return MISFIT_OK;
}

// ---------------------------------------------------------------- capacity
// Per-batch scratch for batches of up to U updates.  The new buffers are
// allocated in full before any old one is freed: a failed growth leaves the
// context exactly as it was (the old capacity still serves smaller batches).
static int ensure_batch_capacity(misfit_ctx* c, int64_t U) {
    if (U <= c->U_cap && c->P.ctr) return MISFIT_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t cap = std::max<int64_t>(std::max<int64_t>(U, 1 << 14), c->U_cap * 2);
    std::vector<void*> A;       // new device buffers
    std::vector<void*> pinned;  // new pinned host buffers
    BP N = c->P;                // graph-level fields stay, batch-level ones are replaced
    Slot ns[MF_NSLOT];
    const u64 bt_size = pow2_at_least((u64)(2 * cap));
    // read logs: one segment per wave of the round grid (256-thread workgroups)
    const int64_t nwaves = (int64_t)c->S.blocks * (256 / MF_WAVE);
    const int64_t fa_cap = std::max<int64_t>(1 << 22, 16 * cap);
    // a growth past the first capacity fails outright under fault injection
    hipError_t e = c->fail_growth && c->U_cap > 0 ? hipErrorOutOfMemory : hipSuccess;
    auto dev = [&](auto** p, int64_t count) {
        if (e == hipSuccess) e = dalloc(A, p, count);
    };
    auto host = [&](auto** p, int64_t bytes) {
        if (e != hipSuccess) return;
        // read by the batch's first kernel over PCIe, and written again for
        // a later batch of the slot: coherent, so no device cache can hold a
        // previous batch's lines
        e = hipHostMalloc((void**)p, (size_t)bytes, hipHostMallocCoherent);
        if (e == hipSuccess) pinned.push_back((void*)*p);
    };
    dev(&N.bt_key, (int64_t)bt_size);
    dev(&N.bt_head, (int64_t)bt_size);
    dev(&N.bt_nxt, cap);
    dev(&N.bt_slot, cap);
    dev(&N.bt_plo, cap);
    dev(&N.bt_phi, cap);
    dev(&N.bt_fl, cap);
    dev(&N.ne_x, 2 * cap);
    dev(&N.ne_y, 2 * cap);
    dev(&N.ne_seq, 2 * cap);
    dev(&N.ne_nxt, 2 * cap);
    dev(&N.seeds, 2 * cap);
    dev(&N.wpot, cap);
    dev(&N.active, cap);
    dev(&N.kind, cap);
    dev(&N.res_w, cap);
    dev(&N.res_x, cap);
    dev(&N.relog, cap);
    dev(&N.rlog, nwaves * MF_RLOG);
    dev(&N.cm, nwaves);
    dev(&N.stage, nwaves);
    dev(&N.fa_z, fa_cap);
    dev(&N.fa_pos, fa_cap);
    for (auto& s : ns) {
        dev(&s.d_u, cap);
        dev(&s.d_v, cap);
        dev(&s.d_ins, cap);
        host(&s.h_u, cap * (int64_t)sizeof(int32_t));
        host(&s.h_v, cap * (int64_t)sizeof(int32_t));
        host(&s.h_ins, cap);
        void* dp = nullptr;
        if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, s.h_u, 0);
        s.g_u = (const int32_t*)dp;
        if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, s.h_v, 0);
        s.g_v = (const int32_t*)dp;
        if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, s.h_ins, 0);
        s.g_ins = (const uint8_t*)dp;
    }
    if (e == hipSuccess) {
        N.bt_mask = bt_size - 1;
        N.fa_cap = fa_cap;
        N.nstage = nwaves;
        mf_fill_i32((int32_t*)N.stage, -1, 4 * nwaves, c->stream);
        mf_fill_u64(N.bt_key, MF_EMPTY_KEY, (int64_t)bt_size, c->stream);
        mf_fill_u32(N.relog, 0u, cap, c->stream);
        mf_fill_i32(N.bt_head, -1, (int64_t)bt_size, c->stream);
        e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        free_list(A);
        for (void* p : pinned) mf_free_host(p);
        const int code = e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? MISFIT_E_NOMEM : MISFIT_E_HIP;
        return fail(c, code, std::string("batch buffers for ") + std::to_string(cap) + " updates: " + hipGetErrorString(e));
    }
    // every allocation succeeded: the old buffers go, the new ones take over
    free_list(c->batch_allocs);
    free_slots(c);
    c->batch_allocs = std::move(A);
    c->P = N;
    for (int i = 0; i < MF_NSLOT; i++) {
        Slot& s = c->slots[i];
        s.h_u = ns[i].h_u; s.h_v = ns[i].h_v; s.h_ins = ns[i].h_ins;
        s.g_u = ns[i].g_u; s.g_v = ns[i].g_v; s.g_ins = ns[i].g_ins;
        s.d_u = ns[i].d_u; s.d_v = ns[i].d_v; s.d_ins = ns[i].d_ins;
    }
    c->bt_size = bt_size;
    c->fa_cap = fa_cap;
    c->U_cap = cap;
    return MISFIT_OK;
}

// Overlay pool and inserted-pair hash sized for `extra` more insertions.
static int ensure_overlay_capacity(misfit_ctx* c, int64_t extra_ins) {
    DevGraph& g = c->P.g;
    int64_t need_ph = 2 * (c->ins_bound + extra_ins) + 1024;
    int64_t need_pool = c->pool_bound + 8 * extra_ins + 16;
    if (need_ph <= c->ph_cap && need_pool <= g.pool_cap) return MISFIT_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (need_pool > g.pool_cap) {
        int64_t nc = std::max<int64_t>(need_pool * 2, g.pool_cap * 2);
        std::vector<void*> fresh;
        int32_t* np = nullptr;
        u64 top = 0;
        hipError_t e = c->fail_growth ? hipErrorOutOfMemory : dalloc(fresh, &np, nc);
        if (e == hipSuccess) e = copy_sync(c, &top, g.pool_top, sizeof(u64), hipMemcpyDeviceToHost);
        if (e == hipSuccess && top) e = copy_sync(c, np, g.pool, top * sizeof(int32_t), hipMemcpyDeviceToDevice);
        if (e != hipSuccess) {  // the old pool stays in use
            free_list(fresh);
            const int code = e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? MISFIT_E_NOMEM : MISFIT_E_HIP;
            return fail(c, code, std::string("overlay pool growth: ") + hipGetErrorString(e));
        }
        c->graph_allocs.push_back(np);
        auto& v = c->graph_allocs;
        auto it = std::find(v.begin(), v.end(), (void*)g.pool);
        if (it != v.end()) { mf_free_device(*it); v.erase(it); }
        g.pool = np;
        g.pool_cap = nc;
    }
    if (need_ph > c->ph_cap) {
        // rehash every key into a table twice as large
        int64_t nc = (int64_t)pow2_at_least((u64)std::max<int64_t>(need_ph * 2, c->ph_cap * 2));
        std::vector<u64> keys((size_t)c->ph_cap);
        std::vector<int32_t> vals((size_t)c->ph_cap);
        HIPCHK(c, copy_sync(c, keys.data(), g.ph_key, c->ph_cap * sizeof(u64), hipMemcpyDeviceToHost));
        HIPCHK(c, copy_sync(c, vals.data(), g.ph_val, c->ph_cap * sizeof(int32_t), hipMemcpyDeviceToHost));
        std::vector<u64> nk((size_t)nc, MF_EMPTY_KEY);
        std::vector<int32_t> nv((size_t)nc, 0);
        auto mixh = [](u64 z) {
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        for (int64_t i = 0; i < c->ph_cap; i++) {
            if (keys[i] == MF_EMPTY_KEY) continue;
            u64 s = mixh(keys[i]) & (u64)(nc - 1);
            while (nk[s] != MF_EMPTY_KEY) s = (s + 1) & (u64)(nc - 1);
            nk[s] = keys[i];
            nv[s] = vals[i];
        }
        // the new table is allocated and filled before the old one goes, so
        // a failure leaves the context on its old, still valid table
        std::vector<void*> fresh;
        u64* nkey = nullptr;
        int32_t* nval = nullptr;
        hipError_t e = c->fail_growth ? hipErrorOutOfMemory : dalloc(fresh, &nkey, nc);
        if (e == hipSuccess) e = dalloc(fresh, &nval, nc);
        if (e == hipSuccess) e = copy_sync(c, nkey, nk.data(), nc * sizeof(u64), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = copy_sync(c, nval, nv.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            free_list(fresh);
            c->S.g = g;
            const int code = e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? MISFIT_E_NOMEM : MISFIT_E_HIP;
            return fail(c, code, std::string("inserted-pair hash growth: ") + hipGetErrorString(e));
        }
        auto& v = c->graph_allocs;
        for (void* old : {(void*)g.ph_key, (void*)g.ph_val}) {
            auto it = std::find(v.begin(), v.end(), old);
            if (it != v.end()) { mf_free_device(*it); v.erase(it); }
        }
        v.push_back(nkey);
        v.push_back(nval);
        g.ph_key = nkey;
        g.ph_val = nval;
        c->ph_cap = nc;
        g.ph_mask = (u64)(nc - 1);
    }
    c->S.g = g;
    return MISFIT_OK;
}

static void reset_overlay(misfit_ctx* c) {
    DevGraph& g = c->P.g;
    hipStream_t st = c->stream;
    mf_fill_i64(g.ov_start, 0, g.n, st);
    mf_fill_i32(g.ov_len, 0, g.n, st);
    mf_fill_i32(g.ov_cap, 0, g.n, st);
    mf_fill_i32(g.ov_alive, 0, g.n, st);
    mf_fill_i32(g.ov_tomb, 0, g.n, st);
    mf_fill_u64(g.pool_top, 0ull, 1, st);
    mf_fill_u64(g.ph_key, MF_EMPTY_KEY, c->ph_cap, st);
    mf_fill_i32(g.ph_val, 0, c->ph_cap, st);
    c->ins_bound = 0;
    c->pool_bound = 0;
}

// ---------------------------------------------------------------- load
extern "C" int misfit_load_csr(misfit_ctx* c, int32_t n, int64_t m, const void* nindex, int nindex_is_64,
                               const int32_t* nlist) {
    if (!c) return MISFIT_E_ARG;
    if (n < 1 || m < 0 || !nindex || (m > 0 && !nlist)) return fail(c, MISFIT_E_ARG, "invalid graph arguments");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the batch scratch keeps per-batch pointers only; graph arrays are new
    free_list(c->graph_allocs);
    free_compact_scratch(c);
    c->has_graph = false;
    c->maximal = false;
    c->U_cap = 0;
    free_list(c->batch_allocs);
    free_slots(c);
    c->P = BP{};
    auto& A = c->graph_allocs;
    DevGraph& g = c->P.g;
    g.n = n;
    g.m = m;
    HIPCHK(c, dalloc(A, &c->d_nindex, (int64_t)n + 1));
    HIPCHK(c, dalloc(A, &c->d_nlist, m));
    c->cap_nindex = ((size_t)n + 1) * 8;
    c->cap_nlist = (size_t)std::max<int64_t>(m, 1) * 4;
    if (nindex_is_64) {
        HIPCHK(c, copy_sync(c, c->d_nindex, nindex, ((size_t)n + 1) * 8, hipMemcpyHostToDevice));
    } else {
        int32_t* tmp = nullptr;
        HIPCHK(c, hipMalloc((void**)&tmp, ((size_t)n + 1) * 4));
        hipError_t e = copy_sync(c, tmp, nindex, ((size_t)n + 1) * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = mf_widen(tmp, c->d_nindex, (int64_t)n + 1, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        mf_free_device(tmp);
        HIPCHK(c, e);
    }
    if (m) HIPCHK(c, copy_sync(c, c->d_nlist, nlist, (size_t)m * 4, hipMemcpyHostToDevice));
    g.nindex = c->d_nindex;
    g.nlist = c->d_nlist;
    HIPCHK(c, dalloc(A, &g.ov_start, n));
    HIPCHK(c, dalloc(A, &g.ov_len, n));
    HIPCHK(c, dalloc(A, &g.ov_cap, n));
    HIPCHK(c, dalloc(A, &g.ov_alive, n));
    HIPCHK(c, dalloc(A, &g.ov_tomb, n));
    g.pool_cap = std::max<int64_t>(1 << 20, m / 16);
    HIPCHK(c, dalloc(A, &g.pool, g.pool_cap));
    HIPCHK(c, dalloc(A, &g.pool_top, 1));
    c->ph_cap = (int64_t)pow2_at_least((u64)std::max<int64_t>(1 << 16, m / 32));
    HIPCHK(c, dalloc(A, &g.ph_key, c->ph_cap));
    HIPCHK(c, dalloc(A, &g.ph_val, c->ph_cap));
    g.ph_mask = (u64)(c->ph_cap - 1);
    HIPCHK(c, dalloc(A, &g.mc, n));
    HIPCHK(c, dalloc(A, &g.seedb, n));
    HIPCHK(c, dalloc(A, &g.wpotb, n));
    HIPCHK(c, dalloc(A, &g.pub, n));
    HIPCHK(c, dalloc(A, &g.evb, n));
    HIPCHK(c, dalloc(A, &g.pfb, n));
    HIPCHK(c, dalloc(A, &g.kw, n));
    HIPCHK(c, dalloc(A, &g.tagw, ((int64_t)n + 31) / 32));
    HIPCHK(c, dalloc(A, &g.nevb, n));
    HIPCHK(c, dalloc(A, &g.cstat, n));
    HIPCHK(c, dalloc(A, &g.mmark, n));
    HIPCHK(c, dalloc(A, &g.cmark, n));
    HIPCHK(c, dalloc(A, &g.cdec, n));
    HIPCHK(c, dalloc(A, &g.ne_head, n));
    HIPCHK(c, dalloc(A, &g.fa_tag, n));
    HIPCHK(c, dalloc(A, &g.fa_n, n));
    HIPCHK(c, dalloc(A, &g.fa_start, n));
    HIPCHK(c, dalloc(A, &g.fa_fill, n));
    BP& P = c->P;
    HIPCHK(c, dalloc(A, &P.pu, n));
    HIPCHK(c, dalloc(A, &P.big, n));
    HIPCHK(c, dalloc(A, &P.cand, n));
    HIPCHK(c, dalloc(A, &P.cdec, n));
    P.pu_cap = n;
    P.big_cap = n;
    HIPCHK(c, dalloc(A, &c->d_ctr, C_NCOUNTERS * MF_CS));
    HIPCHK(c, dalloc(A, &c->d_round, 1));
    HIPCHK(c, dalloc(A, &c->d_acc, 8));
    HIPCHK(c, dalloc(A, &c->d_flags, 2));
    P.ctr = c->d_ctr;
    P.round = c->d_round;
    GraphState& S = c->S;
    HIPCHK(c, dalloc(A, &S.d_cnt, n));
    HIPCHK(c, dalloc(A, &S.d_bmark, n));
    HIPCHK(c, dalloc(A, &S.d_deg, n));
    HIPCHK(c, dalloc(A, &S.d_scan_part, ((int64_t)n + 1) / 2048 + 2));
    S.walk_task_cap = mf_walk_task_need(n, m);
    HIPCHK(c, dalloc(A, (char**)&S.d_walk_tasks, S.walk_task_cap * 16));
    HIPCHK(c, dalloc(A, &S.d_walk_ntasks, 1));
    hipStream_t st = c->stream;
    mf_fill_u32(g.mc, 0u, n, st);
    mf_fill_u32(g.seedb, 0u, n, st);
    mf_fill_u32(g.wpotb, 0u, n, st);
    mf_fill_u32(g.pub, 0u, n, st);
    mf_fill_u32(g.evb, 0u, n, st);
    mf_fill_u32(g.pfb, 0u, n, st);
    mf_fill_u64(g.kw, 0ull, n, st);
    mf_fill_u32(g.tagw, 0u, ((int64_t)n + 31) / 32, st);
    mf_fill_u32(g.nevb, 0u, n, st);
    mf_fill_u32(g.cstat, 0u, n, st);
    mf_fill_u64(g.mmark, 0ull, n, st);
    mf_fill_u64(g.cmark, 0ull, n, st);
    mf_fill_u64(g.cdec, 0ull, n, st);
    mf_fill_i32(g.ne_head, -1, n, st);
    mf_fill_u32(g.fa_tag, 0u, n, st);
    mf_fill_u32(g.fa_n, 0u, n, st);
    mf_fill_i64(g.fa_start, -1, n, st);
    mf_fill_i32(g.fa_fill, 0, n, st);
    mf_fill_u32(S.d_bmark, 0u, n, st);
    mf_fill_u32(c->d_round, 0u, 1, st);
    mf_fill_u64(c->d_ctr, 0ull, C_NCOUNTERS * MF_CS, st);
    // the counters' state at a batch start (each batch's last kernel
    // restores it): zero, no conflict
    mf_fill_u64(c->d_ctr + MF_CS * C_CONF0, ~0ull, 1, st);
    mf_fill_u64(c->d_ctr + MF_CS * C_CONF1, ~0ull, 1, st);
    c->epoch = 0;
    c->bytes_tag = 0;
    S.g = g;
    reset_overlay(c);
    // base validation: monotone offsets first (entry checks index through
    // them), then entry ranges and row order
    HIPCHK(c, mf_validate_csr(S, c->d_flags));
    int flags = 0;
    HIPCHK(c, hipMemcpyAsync(&flags, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (flags & 8) return fail(c, MISFIT_E_ARG, "CSR offsets are not monotone or do not span the entries");
    HIPCHK(c, mf_validate_entries(S, c->d_flags));
    HIPCHK(c, hipMemcpyAsync(&flags, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (flags & 4) return fail(c, MISFIT_E_ARG, "CSR has out-of-range entries");
    g.sorted = (flags & 1) ? 0 : 1;
    S.g = g;
    int rc = build_row_index(c);  // unsorted rows: sorted copy, and no neighbour twice
    if (rc) return rc;
    c->has_graph = true;
    return ensure_batch_capacity(c, 1 << 14);
}

// ---------------------------------------------------------------- MIS
extern "C" int misfit_set_mis(misfit_ctx* c, const uint8_t* mis) {
    if (!c || !mis) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int32_t n = c->P.g.n;
    uint8_t* d = nullptr;
    HIPCHK(c, hipMalloc((void**)&d, n));
    hipError_t e = copy_sync(c, d, mis, n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mf_set_mis(c->S, d);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    mf_free_device(d);
    HIPCHK(c, e);
    // a maximal start stays maximal while the completion runs (a vertex can
    // only lose its last MIS neighbour next to an eviction or at a deleted
    // edge, both in the completion's candidate set), and then the PU scan
    // needs only the vertices the k-count tags
    c->maximal = false;
    HIPCHK(c, mf_check(c->S, c->d_acc));
    HIPCHK(c, hipMemcpyAsync(c->h_small, c->d_acc, 4 * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->h_small[3]) return fail(c, MISFIT_E_CHECK, "edge walk task list overflow");
    c->maximal = c->h_small[1] == 0;
    return MISFIT_OK;
}

extern "C" int misfit_get_mis(misfit_ctx* c, uint8_t* mis) {
    if (!c || !mis) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    int32_t n = c->P.g.n;
    uint8_t* d = nullptr;
    HIPCHK(c, hipMalloc((void**)&d, n));
    hipError_t e = mf_get_mis(c->S, d);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = copy_sync(c, mis, d, n, hipMemcpyDeviceToHost);
    mf_free_device(d);
    HIPCHK(c, e);
    return MISFIT_OK;
}

extern "C" int misfit_flags(misfit_ctx* c, const int32_t* u, const int32_t* v, int64_t count, uint8_t* ins) {
    if (!c || count < 0 || (count && (!u || !v || !ins))) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    if (!count) return MISFIT_OK;
    int32_t n = c->P.g.n;
    for (int64_t i = 0; i < count; i++)
        if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n) return fail(c, MISFIT_E_ARG, "vertex out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    int32_t *du = nullptr, *dv = nullptr;
    uint8_t* di = nullptr;
    hipError_t e = hipMalloc((void**)&du, count * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&dv, count * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&di, count);
    if (e == hipSuccess) e = copy_sync(c, du, u, count * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = copy_sync(c, dv, v, count * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = mf_flags(c->S, du, dv, count, di);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = copy_sync(c, ins, di, count, hipMemcpyDeviceToHost);
    mf_free_device(du);
    mf_free_device(dv);
    mf_free_device(di);
    HIPCHK(c, e);
    return MISFIT_OK;
}

// ---------------------------------------------------------------- flags without allocation
// Staging of misfit_flags_noalloc: the u and v arrays (4 bytes a pair each)
// and the flag bytes of at most MISFIT_NOALLOC_MAX_PAIRS pairs, 9 MiB that
// every GPU the library runs on holds until the process ends.  On the host
// the array's name stands for a host-side stand-in, not device memory, so
// host code never hands it to a kernel or a copy: it uses the address
// hipGetSymbolAddress returned on each device, kept in scratch_ptr.
static constexpr int kScratchDevices = 64;
static constexpr size_t kScratchPairs = (size_t)MISFIT_NOALLOC_MAX_PAIRS;
static __device__ __align__(256) unsigned char scratch[kScratchPairs * 9];
static std::mutex scratch_resolve_mu;
static unsigned char* scratch_ptr[kScratchDevices];
static std::mutex scratch_call_mu[kScratchDevices];  // one call at a time per GPU

// With `device` current: this device's address of the staging array, looked
// up once per device.
static hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_resolve_mu);
    if (scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    scratch_ptr[device] = (unsigned char*)p;
    return hipSuccess;
}

extern "C" int misfit_flags_noalloc(misfit_ctx* c, const int32_t* u, const int32_t* v, int64_t count,
                                    uint8_t* ins) {
    if (!c || count < 0 || (count && (!u || !v || !ins))) return MISFIT_E_ARG;
    if (count > (int64_t)kScratchPairs)
        return fail(c, MISFIT_E_ARG, "misfit_flags_noalloc: " + std::to_string(count) +
                                         " pairs exceed its static staging of MISFIT_NOALLOC_MAX_PAIRS = " +
                                         std::to_string(kScratchPairs) + " (use misfit_flags)");
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    if (!count) return MISFIT_OK;
    int32_t n = c->P.g.n;
    for (int64_t i = 0; i < count; i++)
        if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n) return fail(c, MISFIT_E_ARG, "vertex out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    unsigned char* base = c->device < kScratchDevices ? scratch_ptr[c->device] : nullptr;
    if (!base) return fail(c, MISFIT_E_STATE, "misfit_flags_noalloc: staging array not resolved on this device");
    std::lock_guard<std::mutex> lk(scratch_call_mu[c->device]);
    int32_t* du = (int32_t*)base;
    int32_t* dv = (int32_t*)(base + kScratchPairs * 4);
    uint8_t* di = base + kScratchPairs * 8;
    // every flag byte of the call is written by the kernel, so nothing left
    // by an earlier call needs clearing
    HIPCHK(c, copy_sync(c, du, u, (size_t)count * 4, hipMemcpyHostToDevice));
    HIPCHK(c, copy_sync(c, dv, v, (size_t)count * 4, hipMemcpyHostToDevice));
    HIPCHK(c, mf_flags(c->S, du, dv, count, di));
    HIPCHK(c, copy_sync(c, ins, di, (size_t)count, hipMemcpyDeviceToHost));
    return MISFIT_OK;
}

// ---------------------------------------------------------------- batches
static void fill_stats(const misfit_ctx* c, const Slot& s, misfit_batch_stats* st) {
    const u64* k = s.h_ctr;
    memset(st, 0, sizeof(*st));
    st->updates = s.count;
    st->seeds = (int64_t)k[C_SEEDS];
    st->active_ins = (int64_t)k[C_ACT_INS];
    st->evictions = (int64_t)k[C_EVICT];
    st->admit_ins = (int64_t)k[C_ADMIT_INS];
    st->active_del = (int64_t)k[C_ACT_DEL];
    st->admit_del = (int64_t)k[C_ADMIT_DEL];
    st->completion_adds = (int64_t)k[C_COMP_ADDS];
    st->rounds = (int64_t)k[C_ROUNDS];
    st->evaluations = (int64_t)k[C_SPECS];
    st->maybe_uncovered = (int64_t)k[C_PU];
    st->admissible = (int64_t)k[C_ADM];
    st->delta_ops = (int64_t)k[C_DELTA_OPS];
    st->overlay_entries = (int64_t)k[C_NCOUNTERS];
    st->window = s.P.window;
    // device wall clock (kHz ticks per ms): first kernel start to the end of
    // the tail, and to the start of the first round
    if (c->clock_khz > 0 && k[C_T1] >= k[C_T0]) {
        st->device_start_ms = (double)k[C_T0] / c->clock_khz;
        st->device_end_ms = (double)k[C_T1] / c->clock_khz;
        st->device_ms = (double)(k[C_T1] - k[C_T0]) / c->clock_khz;
        st->prologue_ms = k[C_TP] >= k[C_T0] ? (double)(k[C_TP] - k[C_T0]) / c->clock_khz : st->device_ms;
        if (k[C_TR] >= k[C_TP] && k[C_TR] <= k[C_T1]) {
            st->rounds_ms = (double)(k[C_TR] - k[C_TP]) / c->clock_khz;
            st->counters_ms = k[C_TS] >= k[C_TR] ? (double)(k[C_TS] - k[C_TR]) / c->clock_khz : 0.0;
        }
    }
}

// Host-driven rounds: one round = 4 kernels, then the host reads where the
// committed prefix ends; the completion iterations likewise.
static int drive_plain(misfit_ctx* c, BP& P, const LaunchCfg& L) {
    Range batch("misfit:batch");
    {
        Range pro("misfit:prologue");
        mf_launch_batch_prologue(P, L);
    }
    Range rounds("misfit:rounds");
    for (uint32_t r = 0;; r++) {
        if ((int64_t)r > P.U + 1) return fail(c, MISFIT_E_CHECK, "speculative rounds made no progress");
        int par = (int)(r & 1u);
        if (P.U == 0) break;
        mf_launch_round(P, L, par, r == 0);
        HIPCHK(c, hipMemcpyAsync(c->h_small, P.ctr, C_NCOUNTERS * MF_CS * sizeof(u64), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int64_t start = (int64_t)c->h_small[MF_CS * (C_START0 + (par ^ 1))];
        if (c->debug_rounds) {
            u64 cf = c->h_small[MF_CS * (C_CONF0 + par)];
            fprintf(stderr, "[rounds] batch %u round %u: conflict %lld, next start %lld / %lld\n", P.b, r,
                    cf > (u64)INT64_MAX ? -1LL : (long long)cf, (long long)start, (long long)P.U);
        }
        if (start >= P.U) break;
    }
    Range tail("misfit:tail");
    mf_launch_stats(P, L);
    if (P.completion) {
        mf_launch_completion_collect(P, L);
        for (uint32_t k = 0;; k++) {
            int par = (int)(k & 1u);
            mf_launch_completion_round(P, L, par);
            HIPCHK(c, hipMemcpyAsync(c->h_small, P.ctr + MF_CS * (C_UNDEC0 + par), sizeof(u64), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (c->h_small[0] == 0) break;
            if (k > (uint32_t)P.g.n + 1) return fail(c, MISFIT_E_CHECK, "completion made no progress");
        }
    }
    mf_launch_publish(P, L);
    HIPCHK(c, hipGetLastError());
    return MISFIT_OK;
}

// The queued sequence of one batch: prologue, grid-wide rounds, then the
// one-workgroup tail (leftover rounds, counters, completion, publication).
static void launch_queued(const BP& P, const LaunchCfg& L, int k) {
    Range batch("misfit:batch_enqueue");
    {
        Range pro("misfit:prologue");
        mf_launch_batch_prologue(P, L);
    }
    {
        Range rounds("misfit:rounds");
        for (int r = 0; r < k; r++) mf_launch_round(P, L, r & 1, r == 0);
    }
    Range tail("misfit:tail");
    mf_launch_finish(P, L, k & 1);
}

// The rounds of a batch of `count` updates.  Their number: one round plus
// a high rate of further rounds per update seen on recent batches at least
// a quarter as large (conflicts per update grow with the batch, so larger
// batches bound smaller ones), plus a margin, and at least one round per
// window;
// rounds past the end of the batch cost one empty launch each.
// Their window (edges speculated per round, one per wave of the round
// grid): the whole GPU, or 16 times the edges a round is expected to commit
// when that is less (at least 256), so a batch whose rounds commit few
// edges each leaves the GPU to other graphs.  *sure is false without such
// history (a graph's first batches, or a batch far larger than any before):
// the batch's tail may then halt and have more rounds queued from the host.
struct RoundPlan {
    int k;
    int64_t window;
    bool sure;
};

static RoundPlan plan_rounds(misfit_ctx* c, int64_t count, int64_t full_window) {
    RoundPlan p{0, full_window, true};
    double rpu = 0.0, rpu2 = 0.0;  // largest and second largest eligible rate
    int eligible = 0;
    for (int i = 0; i < 16; i++)
        if (c->hist_count[i] > 0 && 4 * c->hist_count[i] >= count) {
            double r = c->hist_rpu[i];
            if (r > rpu) {
                rpu2 = rpu;
                rpu = r;
            } else if (r > rpu2) {
                rpu2 = r;
            }
            eligible++;
        }
    const bool any = eligible > 0;
    if (c->always_halt) {
        // rounds as expected, not as many as may be needed: a batch that
        // needs more halts and gets them from the host, so no graph's
        // rounds are queued to run empty
        p.sure = count == 0;
        double recent = -1.0;
        for (int j = 1; j <= 16 && recent < 0; j++) {
            int i = (c->hist_pos - j + 32) & 15;
            if (c->hist_count[i] > 0 && 4 * c->hist_count[i] >= count) recent = c->hist_rpu[i];
        }
        if (recent >= 0) {
            double per_round = (double)count / (1.0 + (double)count * recent);
            int64_t w = (int64_t)std::ceil(c->window_factor * per_round);
            p.window = std::min(std::max<int64_t>(256, (w + 31) / 32 * 32), full_window);
        }
        if (c->grid_rounds_fixed) {  // a fixed count; the tail finishes whatever is left
            p.k = c->grid_rounds;
            p.sure = true;
            return p;
        }
        double k = std::ceil((double)count * std::max(recent, 0.0)) + 1.0;
        k = std::max(k, std::ceil((double)count / (double)std::max<int64_t>(p.window, 1)) + 1.0);
        p.k = (int)std::min(k, 65536.0);
        return p;
    }
    if (!any) {
        rpu = 1.0 / 400.0;
        p.sure = count == 0;
    } else {
        double per_round = (double)count / (1.0 + (double)count * rpu);
        int64_t w = (int64_t)std::ceil(c->window_factor * per_round);
        w = std::max<int64_t>(256, (w + 31) / 32 * 32);
        p.window = std::min(w, full_window);
    }
    if (c->grid_rounds_fixed) {
        p.k = c->grid_rounds;
        p.sure = true;
        return p;
    }
    // with enough history the second largest rate (about the 90th
    // percentile of 16), else the largest with 15% more, then a margin that
    // grows with the planned rounds (30% of them, 1 to 2): a round
    // queued past the end costs one empty round of three kernels, a round
    // missing costs a one-workgroup round in the tail, far slower on large
    // batches
    const double base = eligible >= 6 ? std::ceil((double)count * rpu2) : std::ceil((double)count * rpu * 1.15);
    const double mg = c->round_margin >= 0 ? (double)c->round_margin : std::min(2.0, std::max(1.0, std::ceil(0.3 * base)));
    double k = base + mg;
    double per_window = std::ceil((double)count / (double)std::max<int64_t>(p.window, 1)) + 2.0;
    k = std::max(k, per_window);
    p.k = (int)std::min(k, 65536.0);
    return p;
}

// Whole batch enqueued with no host synchronisation inside it.
static int drive_queued(misfit_ctx* c, BP& P, const LaunchCfg& L, int k) {
    launch_queued(P, L, k);
    HIPCHK(c, hipGetLastError());
    return MISFIT_OK;
}

// The tail of slot s's batch halted with edges left (C_HALT): queue more
// grid-wide rounds, as many as the batch's own rate of rounds per committed
// edge predicts for the rest (with a margin unless every batch may halt),
// but at most twice the rounds run so far (at least 8): a rate taken from a
// few rounds is noisy, and doubling bounds the empty rounds to half.  The
// window follows the batch's measured edges per round like a planned one
// (rounds are independent of each other's window).  Then another tail, and
// the batch's end event is recorded again.  Nothing else of this graph
// was queued behind it.
static int continue_batch(misfit_ctx* c, Slot& s) {
    Range range("misfit:continue_rounds");
    const u64* h = s.h_ctr;
    int64_t start = (int64_t)h[C_START0 + s.par_next];
    int64_t left = std::max<int64_t>(s.P.U - start, 0);
    if (h[C_ROUNDS] > 0) {
        if (c->round_threads == 0) {  // the round kernels' size from the batch's own edges per round so far
            LaunchCfg L5 = s.L;
            L5.round_threads = 512;
            L5.round_grid = 0;
            const double per_round = (double)start / (double)h[C_ROUNDS];
            s.L.round_threads =
                1.5 * per_round <= (double)((int64_t)mf_round_grid(L5) * mf_round_wpg(L5)) ? 512 : 256;
        }
        const int64_t wpg = mf_round_wpg(s.L);
        LaunchCfg full = s.L;
        full.round_grid = 0;
        const int64_t full_window = std::max<int64_t>((int64_t)mf_round_grid(full) * wpg, wpg);
        int64_t w = (int64_t)std::ceil(c->window_factor * (double)start / (double)h[C_ROUNDS]);
        w = std::min(std::max<int64_t>(256, (w + 31) / 32 * 32), full_window);
        s.L.round_grid = (int)std::max<int64_t>(w / wpg, 1);
        s.P.window = (int64_t)s.L.round_grid * wpg;
    }
    double per_edge = (double)h[C_ROUNDS] / (double)std::max<int64_t>(start, 1);
    double k = c->always_halt ? std::ceil((double)left * per_edge) + 1.0 : std::ceil((double)left * per_edge * 1.25) + 2.0;
    k = std::min(k, std::max(8.0, 2.0 * (double)h[C_ROUNDS]));
    k = std::max(k, std::ceil((double)left / (double)std::max<int64_t>(s.P.window, 1)));
    int K = (int)std::min(k, 65536.0);
    for (int r = 0; r < K; r++) mf_launch_round(s.P, s.L, (s.par_next + r) & 1);
    s.par_next = (s.par_next + K) & 1;
    mf_launch_finish(s.P, s.L, s.par_next);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(s.done, c->stream));
    return MISFIT_OK;
}

// Blocks until every batch queued without a reliable round estimate has
// finished (queueing further rounds where its tail halted).  Called before
// anything that queues more work on the graph or reads its state.
static int settle(misfit_ctx* c) {
    for (auto& s : c->slots)
        if (s.busy && s.external && !external_done(s))
            return fail(c, MISFIT_E_STATE, "a batch of this graph is still running in the serving kernel");
    for (auto& s : c->slots) {
        if (!s.busy || !s.may_halt) continue;
        while (true) {
            HIPCHK(c, hipEventSynchronize(s.done));
            if (!s.h_ctr[C_HALT]) break;
            int rc = continue_batch(c, s);
            if (rc) return rc;
        }
        s.may_halt = false;
    }
    return MISFIT_OK;
}

// For the serving layer: 1 while a batch of this context may still need
// rounds queued from the host (its graph should get no further batch yet).
int mf_ctx_settling(const misfit_ctx* c) {
    for (const auto& s : c->slots)
        if (s.busy && s.may_halt) return 1;
    return 0;
}

// For the serving layer with one batch per graph on the device: every batch
// may halt, and its rounds are queued as expected rather than as needed.
void mf_ctx_set_halting(misfit_ctx* c, bool on) { c->always_halt = on; }

// Base rows that are not ascending get a sorted copy with positions, so
// adjacency tests binary-search them (any older copy is dropped first).
static int build_row_index(misfit_ctx* c) {
    DevGraph& g = c->P.g;
    if (c->d_sval) mf_free_device(c->d_sval);
    if (c->d_spos) mf_free_device(c->d_spos);
    c->d_sval = c->d_spos = nullptr;
    g.sval = g.spos = nullptr;
    c->S.g = g;
    if (g.sorted || g.m == 0) return MISFIT_OK;
    HIPCHK(c, hipMalloc((void**)&c->d_sval, (size_t)g.m * 4));
    HIPCHK(c, hipMalloc((void**)&c->d_spos, (size_t)g.m * 4));
    c->cap_sval = c->cap_spos = (size_t)g.m * 4;
    mf_fill_i32((int32_t*)c->d_flags, 0, 1, c->stream);
    HIPCHK(c, mf_build_row_index(c->S, c->d_sval, c->d_spos, c->d_flags));
    int flags = 0;
    HIPCHK(c, hipMemcpyAsync(&flags, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flags & 16) return fail(c, MISFIT_E_UNSUPPORTED, "a row lists a neighbour twice");
    if (flags & 32) return fail(c, MISFIT_E_CHECK, "sorted row copy does not match the rows");
    g.sval = c->d_sval;
    g.spos = c->d_spos;
    c->S.g = g;
    return MISFIT_OK;
}

// Row-chunk task list: one task per MF_KCH entries of every row of a
// possibly evicted member, at most U + (base entries + pool slots) / MF_KCH.
static int ensure_ktask_capacity(misfit_ctx* c, int64_t U) {
    int64_t need = U + (c->P.g.m + c->P.g.pool_cap) / MF_KCH + 2;
    if (need <= c->ktask_cap) return MISFIT_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the new list first: on failure the old one stays in use
    const int64_t cap = need + need / 2;
    u64* nt = nullptr;
    const hipError_t e = c->fail_growth && c->ktask_cap > 0 ? hipErrorOutOfMemory
                                                             : hipMalloc((void**)&nt, cap * sizeof(u64));
    if (e != hipSuccess)
        return fail(c, MISFIT_E_NOMEM, std::string("row-chunk task list growth: ") + hipGetErrorString(e));
    if (c->d_ktask) mf_free_device(c->d_ktask);
    c->d_ktask = nt;
    c->ktask_cap = cap;
    return MISFIT_OK;
}

static int check_updates(misfit_ctx* c, const misfit_update* ups, int64_t count, int64_t* nins_out) {
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    if (count > INT32_MAX / 4) return fail(c, MISFIT_E_ARG, "batch too large");
    int32_t n = c->P.g.n;
    int64_t nins = 0;
    for (int64_t i = 0; i < count; i++) {
        if (ups[i].u < 0 || ups[i].u >= n || ups[i].v < 0 || ups[i].v >= n || ups[i].u == ups[i].v)
            return fail(c, MISFIT_E_ARG, "update " + std::to_string(i) + ": endpoints out of range or equal");
        nins += ups[i].is_insertion ? 1 : 0;
    }
    *nins_out = nins;
    return MISFIT_OK;
}

// Queues one batch: from the caller's host array (ups, copied into the
// slot's pinned staging) or from a batch already in device memory (st).
// The window (edges per round) of a batch whose rounds run in one workgroup
// of the serving kernel: MISFIT_SERVE_WINDOW_FACTOR (default 2) times the
// edges the graph's last batch committed per round, rounded up to 32,
// within [64, 1024]; 128 without history.
static int64_t serve_window(const misfit_ctx* c) {
    if (c->serve_window > 0) return c->serve_window;
    if (c->last_epr <= 0.0) return 128;
    const int64_t w = ((int64_t)std::ceil(c->serve_window_factor * c->last_epr) + 31) / 32 * 32;
    return std::min<int64_t>(1024, std::max<int64_t>(64, w));
}

// Queues one batch (ext null), or prepares it for the serving kernel (ext:
// its parameters are returned and nothing is launched; the batch is done
// when the kernel has written its done word, see misfit_poll).
static int submit(misfit_ctx* c, const misfit_update* ups, const misfit_staged* st, int64_t count,
                  int64_t nins, uint64_t* ticket, BP* ext = nullptr, bool ext_whole = false) {
    HIPCHK(c, hipSetDevice(c->device));
    int rc = settle(c);
    if (rc) return rc;
    rc = ensure_batch_capacity(c, count);
    if (rc) return rc;
    rc = ensure_overlay_capacity(c, nins);
    if (rc) return rc;
    rc = ensure_ktask_capacity(c, count);
    if (rc) return rc;
    uint64_t t = c->next_ticket++;
    Slot& s = c->slots[t % MF_NSLOT];
    if (s.busy && s.external) return fail(c, MISFIT_E_STATE, "staging slot still held by a batch in the serving kernel");
    if (s.busy) HIPCHK(c, hipEventSynchronize(s.done));  // staging still owned by an older batch
    if (!st) {
        for (int64_t i = 0; i < count; i++) {
            s.h_u[i] = ups[i].u;
            s.h_v[i] = ups[i].v;
            s.h_ins[i] = ups[i].is_insertion ? 1 : 0;
        }
    }
    s.count = count;
    s.ticket = t;
    s.busy = true;
    c->ins_bound += nins;
    c->pool_bound += 8 * nins;
    c->epoch++;
    BP P = c->P;
    P.g = c->S.g;
    // a batch staged in device memory is read where it lies; one in the
    // slot's pinned staging (not written again before this batch's end
    // event) is read by the first kernel into the slot's device arrays
    P.bu = st ? st->u : s.d_u;
    P.bv = st ? st->v : s.d_v;
    P.bins = st ? st->ins : s.d_ins;
    P.hu = (!count || st) ? nullptr : s.g_u;
    P.hv = (!count || st) ? nullptr : s.g_v;
    P.hins = (!count || st) ? nullptr : s.g_ins;
    P.U = count;
    P.b = c->epoch;
    P.completion = c->completion;
    // sparse scan (tagged vertices only) unless forced dense, the MIS is not
    // known maximal, or the batch may tag more than the dense fraction of n
    const double may_tag = (double)count * (double)P.g.m / (double)std::max<int32_t>(P.g.n, 1);
    const bool dense = c->force_pu_scan || !c->maximal ||
                       (c->pu_dense_fraction >= 0.0 && may_tag > c->pu_dense_fraction * (double)P.g.n);
    P.pu_list = dense ? 0 : 1;
    if (!c->completion) c->maximal = false;
    // grids sized to the batch; one edge per wave of the round grid and round
    LaunchCfg L{c->S.blocks, c->stream, c->round_per_cu};
    L.round_threads = c->round_threads > 0 ? c->round_threads : 256;
    L.U = count;
    L.n = c->P.g.n;
    RoundPlan plan = plan_rounds(c, count, std::max<int64_t>((int64_t)mf_round_grid(L) * mf_round_wpg(L),
                                                             mf_round_wpg(L)));
    if (c->round_threads == 0 && plan.sure) {
        // per batch: 512 threads when the edges a round is expected to
        // commit (the batch over its planned rounds), with half again as
        // margin, fit in the 512-thread grid's waves; without a reliable
        // plan, 256 (a halting batch's continuation decides again)
        LaunchCfg L5 = L;
        L5.round_threads = 512;
        const double per_round = (double)count / (double)std::max(plan.k, 1);
        const int64_t waves5 = (int64_t)mf_round_grid(L5) * mf_round_wpg(L5);
        if (1.5 * per_round <= (double)waves5) {
            L.round_threads = 512;
            plan.window = std::min(plan.window, waves5);  // no more workgroups than the GPU holds
        }
    }
    const int64_t wpg = mf_round_wpg(L);
    L.round_grid = (int)std::max<int64_t>(plan.window / wpg, 1);
    P.window = (int64_t)L.round_grid * wpg;
    P.n_ins = nins;
    P.fin_window = c->fin_window;
    P.ktask = c->d_ktask;
    P.ktask_cap = c->ktask_cap;
    P.h_out = s.h_ctr;
    const bool queued = c->mode == MISFIT_MODE_QUEUED;
    const bool sure = !queued || plan.sure;
    const int k = queued ? plan.k : 0;
    P.may_halt = sure ? 0 : 1;
    c->last_U = count;
    if (ext) {  // the batch's rounds (and with ext_whole its prologue) in the serving kernel
        P.window = P.fin_window = serve_window(c);
        P.may_halt = 0;
        s.P = P;
        s.may_halt = false;
        s.external = true;
        s.h_ctr[MF_DONE_WORD] = 0;
        if (ext_whole) {
            // the kernel runs on another stream: whatever this graph's
            // stream still has queued (buffer fills after a growth or a
            // compaction) must be done before it starts
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            // the prologue's grid-wide kernels on the graph's stream; the
            // item is published once they have ended (s.done)
            mf_launch_batch_prologue(P, L);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(s.done, c->stream));
        *ext = P;
        if (ticket) *ticket = t;
        return MISFIT_OK;
    }
    s.external = false;
    s.P = P;
    s.L = L;
    s.par_next = k & 1;
    s.may_halt = !sure;
    rc = queued ? drive_queued(c, P, L, k) : drive_plain(c, P, L);
    if (rc) return rc;
    // the counters, the pool top and the device clock stamps reached s.h_ctr
    // from the batch's last kernel (the finisher, or the publish kernel of
    // the plain driver)
    HIPCHK(c, hipEventRecord(s.done, c->stream));
    if (ticket) *ticket = t;
    return MISFIT_OK;
}

extern "C" int misfit_submit_batch(misfit_ctx* c, const misfit_update* ups, int64_t count, uint64_t* ticket) {
    if (!c || count < 0 || (count && !ups)) return MISFIT_E_ARG;
    int64_t nins = 0;
    if (int rc = check_updates(c, ups, count, &nins)) return rc;
    return submit(c, ups, nullptr, count, nins, ticket);
}

// For the serving layer: checks and prepares one batch for the serving
// kernel (capacity, staging, parameters); *P receives its parameters.  With
// `whole` nothing is launched: the kernel runs the whole batch.  Otherwise
// the prologue's kernels are queued on the graph's stream, and the item may
// be published once mf_ctx_external_ready reports them ended.  misfit_poll
// reports the batch done once the kernel has written its done word;
// misfit_wait then collects it.
int mf_ctx_prepare_external(misfit_ctx* c, const misfit_update* ups, int64_t count, bool whole, uint64_t* ticket,
                            BP* P) {
    if (!c || count < 0 || (count && !ups) || !P) return MISFIT_E_ARG;
    int64_t nins = 0;
    if (int rc = check_updates(c, ups, count, &nins)) return rc;
    return submit(c, ups, nullptr, count, nins, ticket, P, whole);
}

// 1 once the prologue of a batch prepared for the serving kernel has ended
// (the item may be published), 0 before, or an error code.
int mf_ctx_external_ready(misfit_ctx* c, uint64_t ticket) {
    Slot& s = c->slots[ticket % MF_NSLOT];
    if (!s.busy || s.ticket != ticket || !s.external) return fail(c, MISFIT_E_ARG, "unknown ticket");
    const hipError_t e = hipEventQuery(s.done);
    if (e == hipErrorNotReady) return 0;
    HIPCHK(c, e);
    return 1;
}

extern "C" int misfit_stage_batch(misfit_ctx* c, const misfit_update* ups, int64_t count, misfit_staged** out) {
    if (!c || count < 0 || (count && !ups) || !out) return MISFIT_E_ARG;
    int64_t nins = 0;
    if (int rc = check_updates(c, ups, count, &nins)) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    // one allocation: u, v (int32 each), then the flags
    const size_t nb = (size_t)std::max<int64_t>(count, 1);
    std::vector<uint8_t> h(nb * 9);
    int32_t* hu = (int32_t*)h.data();
    int32_t* hv = hu + nb;
    uint8_t* hi = (uint8_t*)(hv + nb);
    for (int64_t i = 0; i < count; i++) {
        hu[i] = ups[i].u;
        hv[i] = ups[i].v;
        hi[i] = ups[i].is_insertion ? 1 : 0;
    }
    void* d = nullptr;
    HIPCHK(c, hipMalloc(&d, h.size()));
    hipError_t e = copy_sync(c, d, h.data(), h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mf_free_device(d);
        return fail(c, MISFIT_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    misfit_staged* b = new misfit_staged;
    b->u = (int32_t*)d;
    b->v = b->u + nb;
    b->ins = (uint8_t*)(b->v + nb);
    b->count = count;
    b->nins = nins;
    b->n = c->P.g.n;
    c->staged.push_back(b);
    *out = b;
    return MISFIT_OK;
}

static bool owns_staged(const misfit_ctx* c, const misfit_staged* b) {
    return std::find(c->staged.begin(), c->staged.end(), b) != c->staged.end();
}

extern "C" int misfit_submit_staged(misfit_ctx* c, const misfit_staged* b, uint64_t* ticket) {
    if (!c || !b) return MISFIT_E_ARG;
    if (!owns_staged(c, b)) return fail(c, MISFIT_E_ARG, "staged batch does not belong to this context");
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    if (b->n > c->P.g.n) return fail(c, MISFIT_E_STATE, "staged batch was checked against a larger graph");
    return submit(c, nullptr, b, b->count, b->nins, ticket);
}

extern "C" int misfit_release_staged(misfit_ctx* c, misfit_staged* b) {
    if (!c || !b) return MISFIT_E_ARG;
    if (!owns_staged(c, b)) return fail(c, MISFIT_E_ARG, "staged batch does not belong to this context");
    HIPCHK(c, hipSetDevice(c->device));
    // batches that read it may still be queued
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->staged.erase(std::find(c->staged.begin(), c->staged.end(), b));
    mf_free_device(b->u);
    delete b;
    return MISFIT_OK;
}

extern "C" int misfit_poll(misfit_ctx* c, uint64_t ticket) {
    if (!c) return MISFIT_E_ARG;
    Slot& s = c->slots[ticket % MF_NSLOT];
    if (!s.busy || s.ticket != ticket) return fail(c, MISFIT_E_ARG, "unknown or expired ticket");
    if (s.external) return external_done(s) ? 1 : 0;
    hipError_t e = hipEventQuery(s.done);
    if (e == hipErrorNotReady) return 0;
    HIPCHK(c, e);
    if (s.may_halt && s.h_ctr[C_HALT]) {  // halted tail: queue the rest, not done yet
        HIPCHK(c, hipSetDevice(c->device));
        int rc = continue_batch(c, s);
        return rc ? rc : 0;
    }
    return 1;
}

extern "C" int misfit_wait(misfit_ctx* c, uint64_t ticket, misfit_batch_stats* st) {
    if (!c) return MISFIT_E_ARG;
    Slot& s = c->slots[ticket % MF_NSLOT];
    if (!s.busy || s.ticket != ticket) return fail(c, MISFIT_E_ARG, "unknown or expired ticket");
    HIPCHK(c, hipSetDevice(c->device));
    if (s.external) {
        // the serving kernel's caller waits with misfit_poll (it may have to
        // relaunch the kernel); here the batch must be done already
        if (!external_done(s)) return fail(c, MISFIT_E_STATE, "batch still running in the serving kernel");
    } else {
        while (true) {
            HIPCHK(c, hipEventSynchronize(s.done));
            if (!(s.may_halt && s.h_ctr[C_HALT])) break;
            int rc = continue_batch(c, s);
            if (rc) return rc;
        }
    }
    s.busy = false;
    s.may_halt = false;
    const bool was_external = s.external;
    s.external = false;
    if (s.h_ctr[C_ROUNDS] >= 1) c->last_epr = (double)s.count / (double)s.h_ctr[C_ROUNDS];
    if (!was_external && s.count >= 512 && s.h_ctr[C_ROUNDS] >= 1) {
        c->hist_rpu[c->hist_pos & 15] = (double)(s.h_ctr[C_ROUNDS] - 1) / (double)s.count;
        c->hist_count[c->hist_pos & 15] = s.count;
        c->hist_pos++;
    }
    if (s.h_ctr[C_POOL_OVF]) return fail(c, MISFIT_E_CHECK, "overlay pool overflow");
    if (s.h_ctr[C_GUARD]) return fail(c, MISFIT_E_CHECK, "repair loop bound exceeded or task list overflow");
    if (st) fill_stats(c, s, st);
    return MISFIT_OK;
}

extern "C" int misfit_process_batch(misfit_ctx* c, const misfit_update* ups, int64_t count, misfit_batch_stats* st) {
    uint64_t t = 0;
    int rc = misfit_submit_batch(c, ups, count, &t);
    if (rc) return rc;
    return misfit_wait(c, t, st);
}

// ---------------------------------------------------------------- checks
extern "C" int misfit_check(misfit_ctx* c, int64_t* conflicts, int64_t* uncovered, int64_t* count_errors) {
    Range range("misfit:check");
    if (!c) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, mf_check(c->S, c->d_acc));
    HIPCHK(c, hipMemcpyAsync(c->h_small, c->d_acc, 4 * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->h_small[3]) return fail(c, MISFIT_E_CHECK, "edge walk task list overflow");
    if (conflicts) *conflicts = (int64_t)c->h_small[0];
    if (uncovered) *uncovered = (int64_t)c->h_small[1];
    if (count_errors) *count_errors = (int64_t)c->h_small[2];
    return MISFIT_OK;
}

extern "C" int misfit_batch_bytes(misfit_ctx* c, int index_bytes, int64_t* bytes, int64_t* v1, int64_t* v2) {
    if (!c) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    c->bytes_tag++;
    // the last batch's seed count, from its published counters (the device
    // counters are reset when a batch ends)
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->h_small[0] = c->next_ticket > 1 ? c->slots[(c->next_ticket - 1) % MF_NSLOT].h_ctr[C_SEEDS] : 0ull;
    HIPCHK(c, hipMemcpyAsync(c->d_acc + 4, c->h_small, sizeof(u64), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, mf_batch_bytes(c->S, c->P.seeds, c->d_acc + 4, c->bytes_tag, c->d_acc));
    HIPCHK(c, hipMemcpyAsync(c->h_small, c->d_acc, 4 * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t n1 = (int64_t)c->h_small[0], n2 = (int64_t)c->h_small[1], deg = (int64_t)c->h_small[2];
    if (bytes) *bytes = 16 * c->last_U + (int64_t)index_bytes * n1 + 4 * deg + n2;
    if (v1) *v1 = n1;
    if (v2) *v2 = n2;
    return MISFIT_OK;
}

// ---------------------------------------------------------------- compaction
static void free_spares(misfit_ctx* c) {
    for (const misfit_ctx::Spare& s : c->spares) mf_free_device(s.p);
    c->spares.clear();
}

// A kept buffer of at least `bytes` and not much more (best fit), taken out
// of the spares, or null; its capacity in *cap.
static void* take_spare(misfit_ctx* c, size_t bytes, size_t* cap) {
    int best = -1;
    for (size_t i = 0; i < c->spares.size(); i++) {
        const size_t k = c->spares[i].cap;
        if (k >= bytes && k - bytes <= bytes / 4 + 4096 && (best < 0 || k < c->spares[best].cap)) best = (int)i;
    }
    if (best < 0) return nullptr;
    void* p = c->spares[best].p;
    *cap = c->spares[best].cap;
    c->spares.erase(c->spares.begin() + best);
    return p;
}

// A buffer the compaction no longer needs: kept for the next one, or freed.
static void release_buf(misfit_ctx* c, void* p, size_t cap) {
    if (!p) return;
    if (c->keep_compact) c->spares.push_back(misfit_ctx::Spare{p, cap, c->compact_gen});
    else mf_free_device(p);
}

// Whether compactions keep the buffers they release (the serving layer
// decides for its graphs: by default they share one compaction headroom and
// keep none).
void mf_ctx_set_keep_compact(misfit_ctx* c, bool on) {
    c->keep_compact = on;
    if (!on) free_spares(c);
}

// Device bytes the context keeps for its next compaction.
int64_t mf_ctx_spare_bytes(const misfit_ctx* c) {
    int64_t b = 0;
    for (const misfit_ctx::Spare& s : c->spares) b += (int64_t)s.cap;
    return b;
}

// Compaction scratch and kept buffers: freed with the graph (reload, destroy).
static void free_compact_scratch(misfit_ctx* c) {
    free_spares(c);
    for (void* p : c->cs_allocs) mf_free_device(p);
    c->cs_allocs.clear();
    if (c->cs_tmp) mf_free_device(c->cs_tmp);
    c->cs_tmp = nullptr;
    c->cs_tmp_bytes = 0;
    c->cs_n = c->cs_t = 0;
    c->cs_part = c->cs_doff = c->cs_toff = nullptr;
    c->cs_trows = c->cs_big = nullptr;
    c->cs_cnt = nullptr;
    c->cs_tval = c->cs_tidx = c->cs_tsv = c->cs_tsi = nullptr;
}

static hipError_t cs_alloc(misfit_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
    if (e == hipSuccess) c->cs_allocs.push_back(*p);
    return e;
}

// Frees a scratch buffer the context still owns (and nothing else).
static void cs_release(misfit_ctx* c, void* p) {
    if (!p) return;
    auto it = std::find(c->cs_allocs.begin(), c->cs_allocs.end(), p);
    if (it == c->cs_allocs.end()) return;
    c->cs_allocs.erase(it);
    mf_free_device(p);
}

// Per-vertex scratch (once per graph) and the tail lists (grown by half
// again when short).
static hipError_t ensure_compact_scratch(misfit_ctx* c, int64_t ntail) {
    const int64_t n = c->P.g.n;
    hipError_t e = hipSuccess;
    if (c->cs_n != n) {
        free_compact_scratch(c);
        e = cs_alloc(c, (void**)&c->cs_part, (((size_t)n + 2047) / 2048 + 2) * 3 * 8);
        if (e == hipSuccess) e = cs_alloc(c, (void**)&c->cs_doff, ((size_t)n + 1) * 8);
        if (e == hipSuccess) e = cs_alloc(c, (void**)&c->cs_toff, ((size_t)n + 1) * 8);
        if (e == hipSuccess) e = cs_alloc(c, (void**)&c->cs_trows, (size_t)n * 4);
        if (e == hipSuccess) e = cs_alloc(c, (void**)&c->cs_big, (size_t)n * 4);
        if (e == hipSuccess) e = cs_alloc(c, (void**)&c->cs_cnt, 2 * 8);
        if (e != hipSuccess) {
            free_compact_scratch(c);
            return e;
        }
        c->cs_n = n;
    }
    if (ntail > c->cs_t) {
        for (int32_t** p : {&c->cs_tval, &c->cs_tidx, &c->cs_tsv, &c->cs_tsi}) {
            cs_release(c, *p);
            *p = nullptr;
        }
        c->cs_t = 0;
        const int64_t want = ntail + ntail / 2;
        for (int32_t** p : {&c->cs_tval, &c->cs_tidx, &c->cs_tsv, &c->cs_tsi})
            if (e == hipSuccess) e = cs_alloc(c, (void**)p, (size_t)want * 4);
        if (e != hipSuccess) return e;
        c->cs_t = want;
    }
    return hipSuccess;
}

extern "C" int misfit_compact(misfit_ctx* c, int64_t* new_m) {
    if (!c) return MISFIT_E_ARG;
    Range range("misfit:compact");
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    DevGraph& g = c->P.g;
    hipStream_t st = c->stream;
    // every device buffer taken here is released (kept or freed) unless it
    // becomes the new graph; until then the old graph is untouched
    c->compact_gen++;
    std::vector<misfit_ctx::Spare> mine;
    auto grab = [&](void** p, size_t bytes) {
        bytes = std::max<size_t>(bytes, 4);
        size_t cap = 0;
        if (c->keep_compact && (*p = take_spare(c, bytes, &cap))) {
            mine.push_back(misfit_ctx::Spare{*p, cap, 0});
            return hipSuccess;
        }
        cap = c->keep_compact ? bytes + bytes / 32 : bytes;
        hipError_t e = hipMalloc(p, cap);
        if (e != hipSuccess && !c->spares.empty()) {  // out of memory: give the kept buffers back first
            (void)hipGetLastError();
            free_spares(c);
            e = hipMalloc(p, cap);
        }
        if (e == hipSuccess) mine.push_back(misfit_ctx::Spare{*p, cap, 0});
        return e;
    };
    auto drop_all = [&]() {
        for (const misfit_ctx::Spare& b : mine) release_buf(c, b.p, b.cap);
        mine.clear();
    };
    auto keep = [&](void* p) {  // the buffer's capacity
        auto it = std::find_if(mine.begin(), mine.end(), [&](const misfit_ctx::Spare& b) { return b.p == p; });
        const size_t cap = it->cap;
        mine.erase(it);
        return cap;
    };
    int64_t* nidx = nullptr;
    int32_t* nl = nullptr;
    int64_t total = 0, ntail = 0;
    MergeLists L;
    Phases ph;
    ph.to("compact:plan");
    // the new row starts, the deletions and tails before each row and the
    // rows with a tail, in one scan (cs_part: its scratch); the deletion
    // marks of the old rows (position order, and sorted order when a sorted
    // copy exists) and the counting chunks' sums are transient
    hipError_t e = grab((void**)&nidx, ((size_t)g.n + 1) * 8);
    if (e == hipSuccess) e = ensure_compact_scratch(c, 0);
    if (e == hipSuccess)
        e = mf_compact_plan_counts(c->S, nidx, c->cs_doff, c->cs_toff, c->cs_trows, c->cs_cnt, c->cs_part);
    // two read-backs in all, into pinned memory (asynchronous copies): the
    // sizes here, every flag at the end
    ph.to("compact:sizes");
    u64* hs = c->h_small;
    if (e == hipSuccess) e = hipMemcpyAsync(hs, nidx + g.n, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hs + 1, c->cs_toff + g.n, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
        total = (int64_t)hs[0];
        ntail = (int64_t)hs[1];
    }
    ph.to("compact:alloc");
    if (e == hipSuccess) e = ensure_compact_scratch(c, std::max<int64_t>(ntail, 1));  // tail lists never empty
    const int64_t nwords = mf_tomb_words(g.m);
    L.ncst = mf_path_chunks(g.n, g.m);
    // counting chunks: per MF_TCHUNK entries (position order) or per merge-path chunk (sorted order)
    const int64_t nchunks = std::max(mf_tomb_chunks(g.m), L.ncst);
    if (e == hipSuccess) e = grab((void**)&L.tw, (size_t)nwords * sizeof(TombWord));
    if (e == hipSuccess && g.sval) e = grab((void**)&L.tws, (size_t)nwords * sizeof(TombWord));
    if (e == hipSuccess) e = grab((void**)&L.ccnt, (size_t)nchunks * 8);
    if (e == hipSuccess) e = grab((void**)&L.cpre, ((size_t)nchunks + 1) * 8);
    if (e == hipSuccess) e = grab((void**)&L.cpart, ((size_t)nchunks / 2048 + 2) * 8);
    if (e == hipSuccess) e = grab((void**)&L.cst, ((size_t)L.ncst + 1) * 8);
    L.doff = c->cs_doff;
    L.tval = c->cs_tval;
    L.tidx = c->cs_tidx;
    L.toff = c->cs_toff;
    L.trows = c->cs_trows;
    L.ntrows = c->cs_cnt;
    L.tsv = c->cs_tsv;
    L.tsi = c->cs_tsi;
    L.big = c->cs_big;
    L.nbig = c->cs_cnt + 1;
    L.tmp = &c->cs_tmp;
    L.tmp_bytes = &c->cs_tmp_bytes;
    if (e == hipSuccess) e = grab((void**)&nl, (size_t)total * 4);
    // The sorted row copy is merged unless the new rows are ascending for
    // sure (they were, and no row has a tail), without waiting for the copy
    // to tell: when every tail happens to continue its row in ascending
    // order, the merged copy is dropped afterwards.  Tail lists too long to
    // merge take the full sort below.
    const bool merge = total > 0 && !(g.sorted && ntail == 0) && ntail <= (int64_t)INT32_MAX;
    int32_t *nsval = nullptr, *nspos = nullptr;
    if (e == hipSuccess && merge) e = grab((void**)&nsval, (size_t)total * 4);
    if (e == hipSuccess && merge) e = grab((void**)&nspos, (size_t)total * 4);
    ph.to("compact:copy");
    // d_flags[0]: 1 when the new rows are not ascending (k_copy_runs tests
    // whether every tail continues its row in ascending order), 16 and 32
    // from the merge's self-check; d_flags[1]: the self-check of the rows
    if (e == hipSuccess) e = mf_compact_copy(c->S, nidx, nl, L, c->d_flags);
    GraphState NS = c->S;
    NS.g.nindex = nidx;
    NS.g.nlist = nl;
    NS.g.m = total;
    if (e == hipSuccess && c->compact_selfcheck) {  // the new rows' offsets, entries and order
        mf_fill_i32((int32_t*)c->d_flags + 1, 0, 1, st);
        e = mf_validate_csr(NS, c->d_flags + 1);
        if (e == hipSuccess) e = mf_validate_entries(NS, c->d_flags + 1);
    }
    ph.to("compact:merge");
    if (e == hipSuccess && merge)
        e = mf_merge_row_index(c->S, L, ntail, nidx, nl, nsval, nspos, c->d_flags, c->compact_selfcheck);
    ph.to("compact:flags");
    if (e == hipSuccess) e = hipMemcpyAsync(hs, c->d_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        drop_all();
        HIPCHK(c, e);
    }
    int flags[2];
    memcpy(flags, hs, sizeof flags);
    const bool sorted = (flags[0] & 1) == 0;
    if (c->compact_selfcheck && ((flags[1] & ~1) || (sorted && (flags[1] & 1)))) {
        drop_all();
        return fail(c, MISFIT_E_CHECK, "compacted rows failed the self-check (flags " + std::to_string(flags[1]) + ")");
    }
    const bool merged = merge && !sorted;
    if (merged && (flags[0] & 16)) {
        drop_all();
        return fail(c, MISFIT_E_UNSUPPORTED, "a row lists a neighbour twice");
    }
    if (merged && (flags[0] & 32)) {
        drop_all();
        return fail(c, MISFIT_E_CHECK, "merged sorted row copy does not match the compacted rows");
    }
    // the edge walk's split tasks follow the new base's size (allocated
    // before the swap: on failure the old graph stays as it was)
    void* ntasks = nullptr;
    const int64_t walk_need = mf_walk_task_need(g.n, total);
    if (walk_need > c->S.walk_task_cap) {
        e = grab(&ntasks, (size_t)walk_need * 16);
        if (e != hipSuccess) {
            drop_all();
            return fail(c, MISFIT_E_NOMEM, std::string("edge walk tasks: ") + hipGetErrorString(e));
        }
    }
    // the new graph replaces the old one
    ph.to("compact:release");
    const size_t cap_nidx = keep(nidx);
    const size_t cap_nl = keep(nl);
    if (ntasks) {
        keep(ntasks);
        auto& ga = c->graph_allocs;
        auto it = std::find(ga.begin(), ga.end(), c->S.d_walk_tasks);
        if (it != ga.end()) { mf_free_device(*it); ga.erase(it); }
        ga.push_back(ntasks);
        c->S.d_walk_tasks = ntasks;
        c->S.walk_task_cap = walk_need;
    }
    auto& v = c->graph_allocs;
    for (auto old : {std::make_pair((void*)c->d_nindex, c->cap_nindex), std::make_pair((void*)c->d_nlist, c->cap_nlist)}) {
        auto it = std::find(v.begin(), v.end(), old.first);
        if (it != v.end()) {
            v.erase(it);
            release_buf(c, old.first, old.second);
        }
    }
    v.push_back(nidx);
    v.push_back(nl);
    c->d_nindex = nidx;
    c->d_nlist = nl;
    c->cap_nindex = cap_nidx;
    c->cap_nlist = cap_nl;
    release_buf(c, c->d_sval, c->cap_sval);
    release_buf(c, c->d_spos, c->cap_spos);
    c->d_sval = c->d_spos = nullptr;
    if (merged) {
        c->cap_sval = keep(nsval);
        c->cap_spos = keep(nspos);
        c->d_sval = nsval;
        c->d_spos = nspos;
    }
    drop_all();
    // buffers kept by an earlier compaction and not used by this one
    for (size_t i = 0; i < c->spares.size();) {
        if (c->spares[i].gen < c->compact_gen) {
            mf_free_device(c->spares[i].p);
            c->spares.erase(c->spares.begin() + i);
        } else {
            i++;
        }
    }
    g.nindex = nidx;
    g.nlist = nl;
    g.m = total;
    g.sorted = sorted ? 1 : 0;
    g.sval = c->d_sval;
    g.spos = c->d_spos;
    c->S.g = g;
    ph.to("compact:reset_overlay");
    reset_overlay(c);
    if (!sorted && !merged) {
        int rc = build_row_index(c);
        if (rc) return rc;
    }
    if (new_m) *new_m = total;
    return MISFIT_OK;
}

extern "C" int misfit_graph_size(misfit_ctx* c, int32_t* n, int64_t* base_m, int64_t* live_m) {
    if (!c) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    DevGraph& g = c->P.g;
    if (n) *n = g.n;
    if (base_m) *base_m = g.m;
    if (live_m) {
        HIPCHK(c, hipSetDevice(c->device));
        if (int rc_ = settle(c)) return rc_;
        int64_t* nidx = nullptr;
        HIPCHK(c, hipMalloc((void**)&nidx, ((size_t)g.n + 1) * 8));
        hipError_t e = mf_compact_plan(c->S, nidx);
        int64_t total = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&total, nidx + g.n, 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        mf_free_device(nidx);
        HIPCHK(c, e);
        *live_m = total;
    }
    return MISFIT_OK;
}

extern "C" int misfit_export_csr(misfit_ctx* c, int64_t* nindex, int32_t* nlist) {
    if (!c || !nindex) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevGraph& g = c->P.g;
    HIPCHK(c, copy_sync(c, nindex, g.nindex, ((size_t)g.n + 1) * 8, hipMemcpyDeviceToHost));
    if (g.m) {
        if (!nlist) return MISFIT_E_ARG;
        HIPCHK(c, copy_sync(c, nlist, g.nlist, (size_t)g.m * 4, hipMemcpyDeviceToHost));
        // deleted base entries are stored complemented; export live values
        for (int64_t i = 0; i < g.m; i++)
            if (nlist[i] < 0) return fail(c, MISFIT_E_STATE, "base holds deletions: compact before export");
    }
    return MISFIT_OK;
}

extern "C" int misfit_neighbors(misfit_ctx* c, int32_t v, int32_t* out, int64_t cap, int64_t* count) {
    if (!c || !count) return MISFIT_E_ARG;
    if (!c->has_graph) return fail(c, MISFIT_E_STATE, "no graph loaded");
    DevGraph& g = c->P.g;
    if (v < 0 || v >= g.n) return fail(c, MISFIT_E_ARG, "vertex out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if (int rc_ = settle(c)) return rc_;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t be[2];
    int64_t ovs = 0;
    int32_t ovl = 0;
    HIPCHK(c, copy_sync(c, be, g.nindex + v, 16, hipMemcpyDeviceToHost));
    HIPCHK(c, copy_sync(c, &ovl, g.ov_len + v, 4, hipMemcpyDeviceToHost));
    if (ovl) HIPCHK(c, copy_sync(c, &ovs, g.ov_start + v, 8, hipMemcpyDeviceToHost));
    std::vector<int32_t> row((size_t)(be[1] - be[0]) + ovl);
    if (be[1] > be[0]) HIPCHK(c, copy_sync(c, row.data(), g.nlist + be[0], (be[1] - be[0]) * 4, hipMemcpyDeviceToHost));
    if (ovl) HIPCHK(c, copy_sync(c, row.data() + (be[1] - be[0]), g.pool + ovs, (size_t)ovl * 4, hipMemcpyDeviceToHost));
    int64_t k = 0;
    for (int32_t y : row) {
        if (y < 0) continue;
        if (out && k < cap) out[k] = y;
        k++;
    }
    *count = k;
    return MISFIT_OK;
}

extern "C" int misfit_device_info(misfit_ctx* c, char* name, int namelen, int64_t* total_bytes, int64_t* free_bytes) {
    if (!c) return MISFIT_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (name && namelen > 0) {
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
        if (prop.name[0])
            snprintf(name, (size_t)namelen, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        else
            snprintf(name, (size_t)namelen, "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
    }
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    if (total_bytes) *total_bytes = (int64_t)tot;
    if (free_bytes) *free_bytes = (int64_t)fr;
    return MISFIT_OK;
}
