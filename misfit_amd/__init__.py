"""misfit_amd -- MI355X-native incremental maximal-independent-set maintenance.

Python view of the C ABI in include/misfit.h (libmisfit_amd.so, HIP kernels
for gfx950).  The engine runs only on the GPU: if the library is missing or
the device cannot be opened, every entry point raises; there is no CPU
fallback.

Reference interface mirrored (anonymous-artifacts/MISFIT):
  Engine(device)              GPUVertexClusteringGraph ctor  src/misfit/misfit.cu:389-421
  Engine.load_csr / load_egr  FlashGraph::open               core/FlashGraph.h:32-35
  Engine.set_mis / get_mis    setMISFromHost / getMISToHost  src/misfit/misfit.cu:448-462
  Engine.flags                isAdjacent flag at parse time  core/cli_setup.h:66-82
  Engine.process_batch        processBatchGPUVertexClustering src/misfit/misfit.cu:704-730
  Engine.submit / wait        launchBatch / syncAndCollect   src/misfit_async/misfit_async.cu:484-701
  Engine.compact              compact_flash_graph            core/flash_compactor.h:20-105
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmisfit_amd.so")
BIN_DIR = os.path.join(_HERE, "bin")

MODE_PLAIN = 0
MODE_QUEUED = 1
NOALLOC_MAX_PAIRS = 1 << 20  # MISFIT_NOALLOC_MAX_PAIRS

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)

UPDATE_DTYPE = np.dtype([("u", np.int32), ("v", np.int32), ("ins", np.uint8), ("pad", np.uint8, 3)])

ERRORS = {
    -1: "invalid argument", -2: "HIP runtime failure", -3: "call out of order",
    -4: "allocation failed", -5: "unsupported input", -6: "I/O failure", -7: "consistency check failed",
}


class MisfitError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__("%s (%d)%s" % (ERRORS.get(code, "error"), code, (": " + msg) if msg else ""))
        self.code = code


class BatchStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in (
        "updates", "seeds", "active_ins", "evictions", "admit_ins", "active_del", "admit_del",
        "completion_adds", "rounds", "evaluations", "maybe_uncovered", "admissible", "delta_ops",
        "overlay_entries")] + [("device_ms", ctypes.c_double), ("prologue_ms", ctypes.c_double),
                               ("window", ctypes.c_int64), ("device_start_ms", ctypes.c_double),
                               ("device_end_ms", ctypes.c_double), ("rounds_ms", ctypes.c_double),
                               ("counters_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_reserve", ctypes.c_int64), ("compact_fraction", ctypes.c_double),
                ("max_inflight", ctypes.c_int), ("workers", ctypes.c_int), ("persistent", ctypes.c_int)]


class RequestStats(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_int32), ("updates", ctypes.c_int64), ("compactions_before", ctypes.c_int64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("batch", BatchStats)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "batch"}
        d["batch"] = self.batch.as_dict()
        return d


class EgrView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64), ("wide", ctypes.c_int),
                ("nindex", ctypes.c_void_p), ("nlist", ctypes.c_void_p), ("priv", ctypes.c_void_p)]


class EgrReport(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64), ("wide", ctypes.c_int),
                ("bad_offsets", ctypes.c_int64), ("out_of_range", ctypes.c_int64),
                ("self_loops", ctypes.c_int64), ("duplicates", ctypes.c_int64),
                ("unsorted_rows", ctypes.c_int64), ("asymmetric", ctypes.c_int64)]


# every symbol include/misfit.h declares
EXPORTS = (
    "misfit_ctx_create", "misfit_ctx_destroy", "misfit_last_error", "misfit_load_csr",
    "misfit_set_mis", "misfit_get_mis", "misfit_set_rule", "misfit_set_mode", "misfit_flags",
    "misfit_flags_noalloc",
    "misfit_submit_batch", "misfit_wait", "misfit_process_batch", "misfit_sync", "misfit_check",
    "misfit_compact", "misfit_graph_size", "misfit_export_csr", "misfit_neighbors",
    "misfit_batch_bytes", "misfit_device_info", "misfit_egr_open", "misfit_egr_close", "misfit_write_egr",
    "misfit_gen_er", "misfit_gen_rmat", "misfit_gen_rgg", "misfit_greedy_mis", "misfit_gen_updates",
    "misfit_mtx_to_egr", "misfit_egr_check", "misfit_poll", "misfit_timer_begin", "misfit_timer_end",
    "misfit_stage_batch", "misfit_submit_staged", "misfit_release_staged",
    "misfit_server_create", "misfit_server_destroy", "misfit_server_last_error", "misfit_server_add_graph",
    "misfit_server_remove_graph", "misfit_server_submit", "misfit_server_wait", "misfit_server_get_mis",
    "misfit_server_check", "misfit_server_set_reserve",
)

_lib = None


def lib():
    """The loaded libmisfit_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmisfit_amd.so not built: run `make -C misfit_amd/csrc` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.misfit_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.misfit_ctx_destroy.argtypes = [vp]
        L.misfit_last_error.restype = ctypes.c_char_p
        L.misfit_last_error.argtypes = [vp]
        L.misfit_load_csr.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, vp, ctypes.c_int, _i32p]
        L.misfit_set_mis.argtypes = [vp, _u8p]
        L.misfit_get_mis.argtypes = [vp, _u8p]
        L.misfit_set_rule.argtypes = [vp, ctypes.c_int]
        L.misfit_set_mode.argtypes = [vp, ctypes.c_int]
        L.misfit_flags.argtypes = [vp, _i32p, _i32p, ctypes.c_int64, _u8p]
        L.misfit_flags_noalloc.argtypes = [vp, _i32p, _i32p, ctypes.c_int64, _u8p]
        L.misfit_submit_batch.argtypes = [vp, vp, ctypes.c_int64, _u64p]
        L.misfit_stage_batch.argtypes = [vp, vp, ctypes.c_int64, ctypes.POINTER(vp)]
        L.misfit_submit_staged.argtypes = [vp, vp, _u64p]
        L.misfit_release_staged.argtypes = [vp, vp]
        L.misfit_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(BatchStats)]
        L.misfit_process_batch.argtypes = [vp, vp, ctypes.c_int64, ctypes.POINTER(BatchStats)]
        L.misfit_sync.argtypes = [vp]
        L.misfit_timer_begin.argtypes = [vp]
        L.misfit_timer_end.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.misfit_check.argtypes = [vp, _i64p, _i64p, _i64p]
        L.misfit_compact.argtypes = [vp, _i64p]
        L.misfit_graph_size.argtypes = [vp, _i32p, _i64p, _i64p]
        L.misfit_export_csr.argtypes = [vp, _i64p, _i32p]
        L.misfit_neighbors.argtypes = [vp, ctypes.c_int32, _i32p, ctypes.c_int64, _i64p]
        L.misfit_batch_bytes.argtypes = [vp, ctypes.c_int, _i64p, _i64p, _i64p]
        L.misfit_device_info.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, _i64p, _i64p]
        L.misfit_egr_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(EgrView), ctypes.c_char_p, ctypes.c_int]
        L.misfit_egr_close.argtypes = [ctypes.POINTER(EgrView)]
        L.misfit_write_egr.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, _i64p, _i32p,
                                       ctypes.c_char_p, ctypes.c_int]
        L.misfit_poll.argtypes = [vp, ctypes.c_uint64]
        L.misfit_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]
        L.misfit_server_destroy.argtypes = [vp]
        L.misfit_server_last_error.restype = ctypes.c_char_p
        L.misfit_server_last_error.argtypes = [vp]
        L.misfit_server_add_graph.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, vp, ctypes.c_int, _i32p, _u8p,
                                              _i32p]
        L.misfit_server_remove_graph.argtypes = [vp, ctypes.c_int32]
        L.misfit_server_submit.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int64, _u64p]
        L.misfit_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(RequestStats)]
        L.misfit_server_get_mis.argtypes = [vp, ctypes.c_int32, _u8p]
        L.misfit_server_check.argtypes = [vp, ctypes.c_int32, _i64p, _i64p, _i64p]
        L.misfit_server_set_reserve.argtypes = [vp, ctypes.c_int64]
        L.misfit_mtx_to_egr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _i32p, _i64p, ctypes.c_char_p,
                                        ctypes.c_int]
        L.misfit_egr_check.argtypes = [ctypes.c_char_p, ctypes.POINTER(EgrReport), ctypes.c_char_p, ctypes.c_int]
        L.misfit_gen_er.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, _i64p, _i32p, _i64p]
        L.misfit_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, _i64p, _i32p, _i64p]
        L.misfit_gen_rgg.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, _i64p, _i32p, _i64p]
        L.misfit_greedy_mis.argtypes = [ctypes.c_int32, _i64p, _i32p, _u8p]
        L.misfit_gen_updates.argtypes = [ctypes.c_int, ctypes.c_int32, _i64p, _i32p, ctypes.c_int64,
                                         ctypes.c_uint64, _i32p, _i32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _ck(rc, ctx=None):
    if rc != 0:
        msg = ""
        if ctx:
            msg = (lib().misfit_last_error(ctx) or b"").decode(errors="replace")
        raise MisfitError(rc, msg)


# ------------------------------------------------------------------ host I/O
def read_egr(path):
    """(nindex int64[n+1], nlist int32[m], wide) from a .egr file."""
    v = EgrView()
    err = ctypes.create_string_buffer(256)
    rc = lib().misfit_egr_open(os.fsencode(path), ctypes.byref(v), err, 256)
    if rc != 0:
        raise MisfitError(rc, err.value.decode())
    try:
        n, m = v.n, v.m
        if v.wide:
            ni = np.ctypeslib.as_array(ctypes.cast(v.nindex, _i64p), (n + 1,)).astype(np.int64)
        else:
            ni = np.ctypeslib.as_array(ctypes.cast(v.nindex, _i32p), (n + 1,)).astype(np.int64)
        nl = (np.ctypeslib.as_array(ctypes.cast(v.nlist, _i32p), (m,)).copy() if m
              else np.zeros(0, np.int32))
        return ni, nl, bool(v.wide)
    finally:
        lib().misfit_egr_close(ctypes.byref(v))


def write_egr(path, nindex, nlist):
    """Wide .egr, written to a temporary file then renamed over `path`."""
    nindex = np.ascontiguousarray(nindex, np.int64)
    nlist = np.ascontiguousarray(nlist, np.int32)
    err = ctypes.create_string_buffer(256)
    nl = nlist if len(nlist) else np.zeros(1, np.int32)
    rc = lib().misfit_write_egr(os.fsencode(path), len(nindex) - 1, len(nlist), _p(nindex, _i64p),
                                _p(nl, _i32p), err, 256)
    if rc != 0:
        raise MisfitError(rc, err.value.decode())


def write_narrow_egr(path, nindex, nlist):
    """int32-nindex .egr (the reference's narrow form), for workloads that fit."""
    n = len(nindex) - 1
    m = len(nlist)
    assert m < 2 ** 31
    with open(path, "wb") as f:
        f.write(np.array([n, m], np.int32).tobytes())
        f.write(np.asarray(nindex, np.int32).tobytes())
        f.write(np.asarray(nlist, np.int32).tobytes())


def mtx_to_egr(mtx_path, egr_path):
    """Convert a text edge list to the wide .egr form (tools/efficient_converter.py
    semantics); returns (n, m)."""
    n = ctypes.c_int32()
    m = ctypes.c_int64()
    err = ctypes.create_string_buffer(512)
    rc = lib().misfit_mtx_to_egr(os.fsencode(mtx_path), os.fsencode(egr_path), ctypes.byref(n), ctypes.byref(m),
                                 err, 512)
    if rc != 0:
        raise MisfitError(rc, err.value.decode())
    return n.value, m.value


def egr_check(path):
    """Structure report of a .egr file (dict)."""
    r = EgrReport()
    err = ctypes.create_string_buffer(512)
    rc = lib().misfit_egr_check(os.fsencode(path), ctypes.byref(r), err, 512)
    if rc != 0:
        raise MisfitError(rc, err.value.decode())
    return {k: getattr(r, k) for k, _ in EgrReport._fields_}


# ------------------------------------------------------------------ generators
def _csr_from(fn, *args):
    m = ctypes.c_int64()
    _ck(fn(*args, None, None, ctypes.byref(m)))
    return m.value


def gen_er(n, M, seed):
    m = _csr_from(lib().misfit_gen_er, n, M, seed)
    ni = np.zeros(n + 1, np.int64)
    nl = np.zeros(max(m, 1), np.int32)
    mm = ctypes.c_int64()
    _ck(lib().misfit_gen_er(n, M, seed, _p(ni, _i64p), _p(nl, _i32p), ctypes.byref(mm)))
    return ni, nl[:m]


def gen_rmat(scale, edgefactor, seed, a=0.57, b=0.19, c=0.19):
    n = 1 << scale
    ni = np.zeros(n + 1, np.int64)
    # one call: the entries (at most both directions of every generated
    # edge) go to a buffer of that bound, cut to the count after
    nl = np.empty(2 * edgefactor * n, np.int32)
    mm = ctypes.c_int64()
    _ck(lib().misfit_gen_rmat(scale, edgefactor, a, b, c, seed, _p(ni, _i64p), _p(nl, _i32p),
                              ctypes.byref(mm)))
    m = mm.value
    return ni, (nl[:m].copy() if m < len(nl) // 2 else nl[:m])


def gen_rgg(n, avg_degree, seed):
    m = _csr_from(lib().misfit_gen_rgg, n, avg_degree, seed)
    ni = np.zeros(n + 1, np.int64)
    nl = np.zeros(max(m, 1), np.int32)
    mm = ctypes.c_int64()
    _ck(lib().misfit_gen_rgg(n, avg_degree, seed, _p(ni, _i64p), _p(nl, _i32p), ctypes.byref(mm)))
    return ni, nl[:m]


def greedy_mis(nindex, nlist):
    nindex = np.ascontiguousarray(nindex, np.int64)
    nlist = np.ascontiguousarray(nlist, np.int32)
    mis = np.zeros(len(nindex) - 1, np.uint8)
    nl = nlist if len(nlist) else np.zeros(1, np.int32)
    _ck(lib().misfit_greedy_mis(len(nindex) - 1, _p(nindex, _i64p), _p(nl, _i32p), _p(mis, _u8p)))
    return mis


UNIFORM_PAIRS, MIXED_EXISTING, PREFERENTIAL, LOW_DEGREE = 0, 1, 2, 3
# the reference's batch samplers: distinct existing edges, uniform entry /
# vertex by 1/degree / vertex by degree, then a uniform neighbour
EDGE_UNIFORM, EDGE_LOW_DEGREE, EDGE_PREFERENTIAL = 4, 5, 6


def gen_updates(kind, nindex, nlist, count, seed):
    nindex = np.ascontiguousarray(nindex, np.int64)
    nlist = np.ascontiguousarray(nlist, np.int32)
    u = np.zeros(max(count, 1), np.int32)
    v = np.zeros(max(count, 1), np.int32)
    nl = nlist if len(nlist) else np.zeros(1, np.int32)
    _ck(lib().misfit_gen_updates(kind, len(nindex) - 1, _p(nindex, _i64p), _p(nl, _i32p), count, seed,
                                 _p(u, _i32p), _p(v, _i32p)))
    return u[:count], v[:count]


def make_updates(u, v, ins):
    a = np.zeros(len(u), UPDATE_DTYPE)
    a["u"] = u
    a["v"] = v
    a["ins"] = ins
    return a


# ------------------------------------------------------------------ engine
class Engine:
    """One device-resident graph with its overlay and MIS on one GPU."""

    def __init__(self, device=0):
        self.ctx = ctypes.c_void_p()
        _ck(lib().misfit_ctx_create(device, ctypes.byref(self.ctx)))
        self.n = 0

    def close(self):
        if self.ctx:
            lib().misfit_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        _ck(rc, self.ctx)

    def load_csr(self, nindex, nlist):
        nlist = np.ascontiguousarray(nlist, np.int32)
        wide = nindex.dtype == np.int64
        nindex = np.ascontiguousarray(nindex, np.int64 if wide else np.int32)
        n = len(nindex) - 1
        nl = nlist if len(nlist) else np.zeros(1, np.int32)
        self._ck(lib().misfit_load_csr(self.ctx, n, len(nlist), nindex.ctypes.data_as(ctypes.c_void_p),
                                       1 if wide else 0, _p(nl, _i32p)))
        self.n = n

    def load_egr(self, path):
        v = EgrView()
        err = ctypes.create_string_buffer(256)
        rc = lib().misfit_egr_open(os.fsencode(path), ctypes.byref(v), err, 256)
        if rc != 0:
            raise MisfitError(rc, err.value.decode())
        try:
            self._ck(lib().misfit_load_csr(self.ctx, v.n, v.m, v.nindex, v.wide,
                                           ctypes.cast(v.nlist, _i32p)))
            self.n = v.n
        finally:
            lib().misfit_egr_close(ctypes.byref(v))

    def set_mis(self, mis):
        mis = np.ascontiguousarray(mis, np.uint8)
        assert len(mis) == self.n
        self._ck(lib().misfit_set_mis(self.ctx, _p(mis, _u8p)))

    def get_mis(self):
        out = np.zeros(self.n, np.uint8)
        self._ck(lib().misfit_get_mis(self.ctx, _p(out, _u8p)))
        return out

    def set_rule(self, completion=True):
        self._ck(lib().misfit_set_rule(self.ctx, 1 if completion else 0))

    def set_mode(self, mode):
        self._ck(lib().misfit_set_mode(self.ctx, mode))

    def flags(self, u, v):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.zeros(max(len(u), 1), np.uint8)
        if len(u):
            self._ck(lib().misfit_flags(self.ctx, _p(u, _i32p), _p(v, _i32p), len(u), _p(ins, _u8p)))
        return ins[: len(u)]

    def flags_noalloc(self, u, v):
        """As flags(), staged in the library's static device array: no
        allocation, at most NOALLOC_MAX_PAIRS pairs."""
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.zeros(max(len(u), 1), np.uint8)
        if len(u):
            self._ck(lib().misfit_flags_noalloc(self.ctx, _p(u, _i32p), _p(v, _i32p), len(u), _p(ins, _u8p)))
        return ins[: len(u)]

    def submit(self, updates):
        updates = np.ascontiguousarray(updates, UPDATE_DTYPE)
        t = ctypes.c_uint64()
        self._ck(lib().misfit_submit_batch(self.ctx, updates.ctypes.data_as(ctypes.c_void_p), len(updates),
                                           ctypes.byref(t)))
        return t.value

    def stage(self, updates):
        """Checks a batch and uploads it to device memory; returns a handle
        for submit_staged (freed by release_staged or with the engine)."""
        updates = np.ascontiguousarray(updates, UPDATE_DTYPE)
        h = ctypes.c_void_p()
        self._ck(lib().misfit_stage_batch(self.ctx, updates.ctypes.data_as(ctypes.c_void_p), len(updates),
                                          ctypes.byref(h)))
        return h

    def submit_staged(self, handle):
        """Queues a staged batch (no host copy); returns its ticket."""
        t = ctypes.c_uint64()
        self._ck(lib().misfit_submit_staged(self.ctx, handle, ctypes.byref(t)))
        return t.value

    def release_staged(self, handle):
        """Waits for the engine's stream and frees a staged batch."""
        self._ck(lib().misfit_release_staged(self.ctx, handle))

    def wait(self, ticket):
        st = BatchStats()
        self._ck(lib().misfit_wait(self.ctx, ticket, ctypes.byref(st)))
        return st.as_dict()

    def poll(self, ticket):
        """True once the batch has finished on the device (never blocks)."""
        rc = lib().misfit_poll(self.ctx, ticket)
        if rc < 0:
            self._ck(rc)
        return rc == 1

    def process_batch(self, updates):
        updates = np.ascontiguousarray(updates, UPDATE_DTYPE)
        st = BatchStats()
        self._ck(lib().misfit_process_batch(self.ctx, updates.ctypes.data_as(ctypes.c_void_p), len(updates),
                                            ctypes.byref(st)))
        return st.as_dict()

    def sync(self):
        self._ck(lib().misfit_sync(self.ctx))

    def timer_begin(self):
        """Records a HIP event where the engine's stream stands now."""
        self._ck(lib().misfit_timer_begin(self.ctx))

    def timer_end(self):
        """Records a second event after all queued work, waits for it and
        returns the milliseconds since timer_begin's event."""
        ms = ctypes.c_double()
        self._ck(lib().misfit_timer_end(self.ctx, ctypes.byref(ms)))
        return ms.value

    def check(self):
        c, u, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._ck(lib().misfit_check(self.ctx, ctypes.byref(c), ctypes.byref(u), ctypes.byref(e)))
        return c.value, u.value, e.value

    def compact(self):
        m = ctypes.c_int64()
        self._ck(lib().misfit_compact(self.ctx, ctypes.byref(m)))
        return m.value

    def graph_size(self):
        n, bm, lm = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        self._ck(lib().misfit_graph_size(self.ctx, ctypes.byref(n), ctypes.byref(bm), ctypes.byref(lm)))
        return n.value, bm.value, lm.value

    def export_csr(self):
        n, bm, _ = self.graph_size()
        ni = np.zeros(n + 1, np.int64)
        nl = np.zeros(max(bm, 1), np.int32)
        self._ck(lib().misfit_export_csr(self.ctx, _p(ni, _i64p), _p(nl, _i32p)))
        return ni, nl[:bm]

    def neighbors(self, v):
        k = ctypes.c_int64()
        self._ck(lib().misfit_neighbors(self.ctx, v, None, 0, ctypes.byref(k)))
        out = np.zeros(max(k.value, 1), np.int32)
        self._ck(lib().misfit_neighbors(self.ctx, v, _p(out, _i32p), k.value, ctypes.byref(k)))
        return out[: k.value]

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        tot, fr = ctypes.c_int64(), ctypes.c_int64()
        self._ck(lib().misfit_device_info(self.ctx, name, 256, ctypes.byref(tot), ctypes.byref(fr)))
        return name.value.decode(), tot.value, fr.value

    def batch_bytes(self, index_bytes=8):
        b, v1, v2 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._ck(lib().misfit_batch_bytes(self.ctx, index_bytes, ctypes.byref(b), ctypes.byref(v1),
                                          ctypes.byref(v2)))
        return b.value, v1.value, v2.value


class Server:
    """Serving layer (misfit_server_*): many graphs on one GPU, update batches
    submitted as they arrive, continuous batching across graphs."""

    def __init__(self, device=0, hbm_reserve=0, compact_fraction=0.05, max_inflight=2, workers=0, persistent=0):
        self.srv = ctypes.c_void_p()
        opts = ServerOpts(hbm_reserve, compact_fraction, max_inflight, workers, persistent)
        _ck(lib().misfit_server_create(device, ctypes.byref(opts), ctypes.byref(self.srv)))

    def _ck(self, rc):
        if rc != 0:
            raise MisfitError(rc, (lib().misfit_server_last_error(self.srv) or b"").decode(errors="replace"))

    def close(self):
        if self.srv:
            lib().misfit_server_destroy(self.srv)
            self.srv = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_graph(self, nindex, nlist, mis):
        nindex = np.ascontiguousarray(nindex, np.int64)
        nlist = np.ascontiguousarray(nlist, np.int32)
        mis = np.ascontiguousarray(mis, np.uint8)
        gid = ctypes.c_int32()
        nl = nlist if len(nlist) else np.zeros(1, np.int32)
        self._ck(lib().misfit_server_add_graph(self.srv, len(nindex) - 1, len(nlist), nindex.ctypes.data, 1,
                                               _p(nl, _i32p), _p(mis, _u8p), ctypes.byref(gid)))
        return gid.value

    def remove_graph(self, gid):
        self._ck(lib().misfit_server_remove_graph(self.srv, gid))

    def set_reserve(self, hbm_reserve):
        """Device bytes kept free for every later admission (graphs, batches, compactions)."""
        self._ck(lib().misfit_server_set_reserve(self.srv, int(hbm_reserve)))

    def submit(self, gid, updates):
        updates = np.ascontiguousarray(updates, UPDATE_DTYPE)
        t = ctypes.c_uint64()
        self._ck(lib().misfit_server_submit(self.srv, gid, updates.ctypes.data, len(updates), ctypes.byref(t)))
        return t.value

    def wait(self, request):
        """Request statistics (dict); raises MisfitError on a refused or failed batch."""
        st = RequestStats()
        self._ck(lib().misfit_server_wait(self.srv, request, ctypes.byref(st)))
        return st.as_dict()

    def get_mis(self, gid, n):
        out = np.zeros(n, np.uint8)
        self._ck(lib().misfit_server_get_mis(self.srv, gid, _p(out, _u8p)))
        return out

    def check(self, gid):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._ck(lib().misfit_server_check(self.srv, gid, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value
