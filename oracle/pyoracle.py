"""ctypes view of the CPU oracle (oracle/misfit_oracle.c) and of the reference's
own host classes built into oracle/_ref/ (oracle/ref_driver.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
benchmark's CPU-baseline leg.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libmisfit_ref.so")

COMPLETION = 1
LITERAL = 2

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Stats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in (
        "active_ins", "evictions", "admit_ins", "active_del", "admit_del",
        "completion_adds", "seeds", "rows", "row_elems")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    """Compile liboracle.so (and _ref when the reference checkout is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp = ctypes.c_void_p
        L.or_load_egr.restype = vp
        L.or_load_egr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.or_from_csr.restype = vp
        L.or_from_csr.argtypes = [ctypes.c_int32, ctypes.c_int64, _i64p, _i32p]
        L.or_free.argtypes = [vp]
        L.or_nodes.restype = ctypes.c_int32
        L.or_nodes.argtypes = [vp]
        L.or_base_edges.restype = ctypes.c_int64
        L.or_base_edges.argtypes = [vp]
        L.or_is_wide.argtypes = [vp]
        L.or_delta_ops.restype = ctypes.c_int64
        L.or_delta_ops.argtypes = [vp]
        L.or_base_csr.argtypes = [vp, _i64p, _i32p]
        L.or_is_adjacent.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.or_insert_edge.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.or_remove_edge.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.or_neighbors.restype = ctypes.c_int64
        L.or_neighbors.argtypes = [vp, ctypes.c_int32, _i32p, ctypes.c_int64]
        L.or_degree.restype = ctypes.c_int64
        L.or_degree.argtypes = [vp, ctypes.c_int32]
        L.or_flags.argtypes = [vp, _i32p, _i32p, ctypes.c_int64, _u8p]
        L.or_apply_batch.argtypes = [vp, _i32p, _i32p, _u8p, ctypes.c_int64]
        L.or_repair.argtypes = [vp, _u8p, _i32p, _i32p, _u8p, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(Stats)]
        L.or_process_batch.argtypes = L.or_repair.argtypes
        L.or_check.argtypes = [vp, _u8p, _i64p, _i64p]
        L.or_compact.restype = ctypes.c_int64
        L.or_compact.argtypes = [vp, _i64p, _i32p]
        L.or_greedy_mis.argtypes = [vp, _u8p]
        L.or_parse_batch_file.restype = ctypes.c_int64
        L.or_parse_batch_file.argtypes = [ctypes.c_char_p, vp, _i32p, _i32p, _u8p, ctypes.c_int64]
        L.or_read_mis.argtypes = [ctypes.c_char_p, ctypes.c_int32, _u8p, _i64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Graph:
    """Base CSR plus the reference-order delta overlay, on the CPU."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load_egr(cls, path):
        err = ctypes.create_string_buffer(256)
        h = lib().or_load_egr(os.fsencode(path), err, 256)
        if not h:
            raise IOError(err.value.decode())
        return cls(h)

    @classmethod
    def from_csr(cls, nindex, nlist):
        nindex = np.ascontiguousarray(nindex, dtype=np.int64)
        nlist = np.ascontiguousarray(nlist, dtype=np.int32)
        n = len(nindex) - 1
        if len(nlist) == 0:
            nlist = np.zeros(1, np.int32)
            m = 0
        else:
            m = len(nlist)
        return cls(lib().or_from_csr(n, m, _p(nindex, _i64p), _p(nlist, _i32p)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().or_nodes(self.h)

    @property
    def m(self):
        return lib().or_base_edges(self.h)

    @property
    def wide(self):
        return bool(lib().or_is_wide(self.h))

    @property
    def delta_ops(self):
        return lib().or_delta_ops(self.h)

    def base_csr(self):
        ni = np.zeros(self.n + 1, np.int64)
        nl = np.zeros(max(self.m, 1), np.int32)
        lib().or_base_csr(self.h, _p(ni, _i64p), _p(nl, _i32p))
        return ni, nl[: self.m]

    def is_adjacent(self, u, v):
        return bool(lib().or_is_adjacent(self.h, u, v))

    def insert_edge(self, u, v):
        lib().or_insert_edge(self.h, u, v)

    def remove_edge(self, u, v):
        lib().or_remove_edge(self.h, u, v)

    def neighbors(self, v):
        k = lib().or_neighbors(self.h, v, None, 0)
        out = np.zeros(max(k, 1), np.int32)
        lib().or_neighbors(self.h, v, _p(out, _i32p), k)
        return out[:k]

    def degree(self, v):
        return lib().or_degree(self.h, v)

    def flags(self, u, v):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.zeros(len(u), np.uint8)
        lib().or_flags(self.h, _p(u, _i32p), _p(v, _i32p), len(u), _p(ins, _u8p))
        return ins

    def apply_batch(self, u, v, ins):
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.ascontiguousarray(ins, np.uint8)
        lib().or_apply_batch(self.h, _p(u, _i32p), _p(v, _i32p), _p(ins, _u8p), len(u))

    def repair(self, mis, u, v, ins, completion=True, literal=False, threads=1, apply=False):
        """Repairs `mis` (uint8 array, modified in place) for one batch."""
        assert mis.dtype == np.uint8 and mis.flags.c_contiguous
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.ascontiguousarray(ins, np.uint8)
        mode = (COMPLETION if completion else 0) | (LITERAL if literal else 0)
        st = Stats()
        fn = lib().or_process_batch if apply else lib().or_repair
        fn(self.h, _p(mis, _u8p), _p(u, _i32p), _p(v, _i32p), _p(ins, _u8p), len(u), mode,
           threads, ctypes.byref(st))
        return st.as_dict()

    def process_batch(self, mis, u, v, ins, **kw):
        return self.repair(mis, u, v, ins, apply=True, **kw)

    def check(self, mis):
        c = ctypes.c_int64()
        un = ctypes.c_int64()
        lib().or_check(self.h, _p(mis, _u8p), ctypes.byref(c), ctypes.byref(un))
        return c.value // 2, un.value

    def compact(self):
        m = lib().or_compact(self.h, None, None)
        ni = np.zeros(self.n + 1, np.int64)
        nl = np.zeros(max(m, 1), np.int32)
        lib().or_compact(self.h, _p(ni, _i64p), _p(nl, _i32p))
        return ni, nl[:m]

    def greedy_mis(self):
        mis = np.zeros(self.n, np.uint8)
        lib().or_greedy_mis(self.h, _p(mis, _u8p))
        return mis

    def parse_batch_file(self, path):
        k = lib().or_parse_batch_file(os.fsencode(path), self.h, None, None, None, 0)
        if k < 0:
            raise IOError(path)
        u = np.zeros(max(k, 1), np.int32)
        v = np.zeros(max(k, 1), np.int32)
        ins = np.zeros(max(k, 1), np.uint8)
        lib().or_parse_batch_file(os.fsencode(path), self.h, _p(u, _i32p), _p(v, _i32p),
                                  _p(ins, _u8p), k)
        return u[:k], v[:k], ins[:k]


def read_mis(path, n):
    mis = np.zeros(n, np.uint8)
    card = ctypes.c_int64()
    rc = lib().or_read_mis(os.fsencode(path), n, _p(mis, _u8p), ctypes.byref(card))
    if rc != 0:
        raise ValueError("unparsable line in %s" % path)
    return mis, card.value


def batch_files_sorted(folder):
    """Every regular file, ordered by the integer value of its stem
    (core/cli_setup.h:49-59)."""
    paths = [os.path.join(folder, f) for f in os.listdir(folder)
             if os.path.isfile(os.path.join(folder, f))]
    return sorted(paths, key=lambda p: int(os.path.splitext(os.path.basename(p))[0]))


# ------------------------------------------------------------ reference build
_ref = None


def ref_available():
    return os.path.exists(_REF)


def reflib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(_REF)
        vp = ctypes.c_void_p
        L.ref_open.restype = vp
        L.ref_open.argtypes = [ctypes.c_char_p]
        L.ref_close.argtypes = [vp]
        L.ref_nodes.restype = ctypes.c_int32
        L.ref_nodes.argtypes = [vp]
        L.ref_base_edges.restype = ctypes.c_int64
        L.ref_base_edges.argtypes = [vp]
        L.ref_base_nindex.restype = ctypes.c_int64
        L.ref_base_nindex.argtypes = [vp, ctypes.c_int32]
        L.ref_is_adjacent.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.ref_insert.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.ref_remove.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.ref_degree.restype = ctypes.c_int64
        L.ref_degree.argtypes = [vp, ctypes.c_int32]
        L.ref_delta_ops.restype = ctypes.c_int64
        L.ref_delta_ops.argtypes = [vp]
        L.ref_neighbors.restype = ctypes.c_int64
        L.ref_neighbors.argtypes = [vp, ctypes.c_int32, _i32p, ctypes.c_int64]
        L.ref_compact.argtypes = [vp, ctypes.c_char_p]
        L.ref_apply.argtypes = [vp, _i32p, _i32p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int64, ctypes.c_int]
        _ref = L
    return _ref


class RefGraph:
    """The reference's HybridGraph (core/FlashGraph.h:278-311) via oracle/_ref."""

    def __init__(self, path):
        self.h = reflib().ref_open(os.fsencode(path))

    def close(self):
        if self.h:
            reflib().ref_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def n(self):
        return reflib().ref_nodes(self.h)

    @property
    def m(self):
        return reflib().ref_base_edges(self.h)

    def is_adjacent(self, u, v):
        return bool(reflib().ref_is_adjacent(self.h, u, v))

    def insert_edge(self, u, v):
        reflib().ref_insert(self.h, u, v)

    def remove_edge(self, u, v):
        reflib().ref_remove(self.h, u, v)

    def degree(self, v):
        return reflib().ref_degree(self.h, v)

    @property
    def delta_ops(self):
        return reflib().ref_delta_ops(self.h)

    def neighbors(self, v):
        k = reflib().ref_neighbors(self.h, v, None, 0)
        out = np.zeros(max(k, 1), np.int32)
        reflib().ref_neighbors(self.h, v, _p(out, _i32p), k)
        return out[:k]

    def nindex(self, v):
        return reflib().ref_base_nindex(self.h, v)

    def compact_to(self, path):
        reflib().ref_compact(self.h, os.fsencode(path))

    def apply(self, u, v, ins, check=True):
        """Applies a batch to the overlay the way src/misfit/misfit.cu:708-714 does
        (check=False skips the adjacency tests: the same overlay when every
        update changes the graph, each pair once)."""
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        ins = np.ascontiguousarray(ins, np.uint8)
        reflib().ref_apply(self.h, _p(u, _i32p), _p(v, _i32p), ins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                           len(u), 1 if check else 0)
