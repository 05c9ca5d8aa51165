"""CPU-side checks of libmisfit_amd.so: it loads, exports every entry point
include/misfit.h declares, and its host parts (.egr I/O, generators, greedy
MIS, update streams, the drop-in binaries' argument handling) behave as
specified.  No call here touches a GPU."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from tests import helpers as H
from tests.conftest import ROOT, SAMPLE


def _declared():
    hdr = open(os.path.join(ROOT, "include", "misfit.h")).read()
    return sorted(set(re.findall(r"\b(misfit_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol(engine_mod):
    L = engine_mod.lib()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert sorted(engine_mod.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", engine_mod.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    defined = set(re.findall(r" T (misfit_[a-z0-9_]+)", out))
    assert set(names) <= defined


def test_library_has_gfx950_code(engine_mod):
    blob = open(engine_mod.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_egr_read_both_forms(engine_mod, oracle, tmp_path):
    ni, nl, wide = engine_mod.read_egr(os.path.join(SAMPLE, "sample.egr"))
    g = oracle.Graph.load_egr(os.path.join(SAMPLE, "sample.egr"))
    oni, onl = g.base_csr()
    assert wide and np.array_equal(ni, oni) and np.array_equal(nl, onl)
    p = str(tmp_path / "narrow.egr")
    engine_mod.write_narrow_egr(p, ni, nl)
    raw = open(p, "rb").read()
    assert struct.unpack("<ii", raw[:8]) == (2000, 16000)
    ni2, nl2, wide2 = engine_mod.read_egr(p)
    assert not wide2 and np.array_equal(ni2, ni) and np.array_equal(nl2, nl)
    g2 = oracle.Graph.load_egr(p)
    assert not g2.wide and np.array_equal(g2.base_csr()[1], nl)


def test_egr_writer_layout_and_atomic_replace(engine_mod, tmp_path):
    rng = np.random.default_rng(1)
    ni, nl = H.random_graph(rng, 50, 0.1)
    p = tmp_path / "g.egr"
    p.write_bytes(b"old contents")
    engine_mod.write_egr(str(p), ni, nl)
    raw = p.read_bytes()
    n, sent, m = struct.unpack("<iiq", raw[:16])
    assert (n, sent, m) == (50, -1, len(nl))
    assert raw[16:16 + 51 * 8] == ni.astype(np.int64).tobytes()
    assert raw[16 + 51 * 8:] == nl.astype(np.int32).tobytes()
    assert [f for f in os.listdir(tmp_path) if ".tmp." in f] == []


@pytest.mark.parametrize("content,err", [
    (struct.pack("<ii", 0, 5), "invalid header"),
    (struct.pack("<ii", 3, -2), "invalid header"),
    (struct.pack("<ii", 3, 4) + b"\0" * 8, "truncated"),
    (b"\1\0", "header"),
])
def test_egr_reader_rejects_bad_files(engine_mod, tmp_path, content, err):
    p = tmp_path / "bad.egr"
    p.write_bytes(content)
    with pytest.raises(engine_mod.MisfitError) as ei:
        engine_mod.read_egr(str(p))
    assert err in str(ei.value)


def _csr_ok(ni, nl):
    n = len(ni) - 1
    src = np.repeat(np.arange(n), np.diff(ni))
    assert not np.any(src == nl), "self-loop"
    for x in range(0, n, max(1, n // 500)):
        r = nl[ni[x]:ni[x + 1]]
        assert np.all(np.diff(r) > 0), "row not strictly ascending"
    key = src.astype(np.int64) * n + nl
    rkey = nl.astype(np.int64) * n + src
    assert np.array_equal(np.sort(key), np.sort(rkey)), "not symmetric"


def test_generators_valid_and_deterministic(engine_mod):
    M = engine_mod
    ni, nl = M.gen_er(5000, 40000, 42)
    _csr_ok(ni, nl)
    assert len(nl) == 80000
    a, b = M.gen_er(5000, 40000, 42)
    assert np.array_equal(a, ni) and np.array_equal(b, nl)
    ri, rl = M.gen_rmat(12, 16, 7)
    _csr_ok(ri, rl)
    assert len(ri) == 4097
    deg = np.diff(ri)
    assert deg[:16].mean() > 10 * deg.mean()  # skew towards low ids (a = 0.57)
    gi, gl = M.gen_rgg(20000, 32.0, 11)
    _csr_ok(gi, gl)
    assert 28 < len(gl) / 20000 < 34  # boundary effects trim the mean a little
    gi2, gl2 = M.gen_rgg(20000, 32.0, 11)
    assert np.array_equal(gl, gl2)


def test_greedy_mis_matches_oracle(engine_mod, oracle):
    rng = np.random.default_rng(2)
    for _ in range(3):
        ni, nl = H.random_graph(rng, 300, 0.03)
        assert np.array_equal(engine_mod.greedy_mis(ni, nl), oracle.Graph.from_csr(ni, nl).greedy_mis())
    ni, nl, _ = engine_mod.read_egr(os.path.join(SAMPLE, "sample.egr"))
    mis, _ = oracle.read_mis(os.path.join(SAMPLE, "sample_mis.txt"), 2000)
    assert np.array_equal(engine_mod.greedy_mis(ni, nl), mis)


def test_update_streams(engine_mod, oracle):
    M = engine_mod
    ni, nl = M.gen_rmat(12, 16, 7)
    g = oracle.Graph.from_csr(ni, nl)
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, 4000, 7)
    ins = g.flags(u, v)
    assert ins.sum() == 2000  # half insertions of non-edges, half deletions
    pairs = {(min(a, b), max(a, b)) for a, b in zip(u.tolist(), v.tolist())}
    assert len(pairs) == 4000  # distinct across the stream
    u0, v0 = M.gen_updates(M.UNIFORM_PAIRS, ni, nl, 5000, 1)
    assert np.all(u0 != v0) and u0.min() >= 0 and u0.max() < 4096
    u2, v2 = M.gen_updates(M.PREFERENTIAL, ni, nl, 5000, 1)
    deg = np.diff(ni)
    assert deg[u2].mean() > 2 * deg.mean()
    u3, v3 = M.gen_updates(M.LOW_DEGREE, ni, nl, 5000, 1)
    q = np.sort(deg)[len(deg) // 4 - 1]
    assert deg[u3].max() <= q and deg[v3].max() <= q


def test_bench_refuses_more_than_one_gpu():
    """The bench measures one GPU: --gpus above 1 exits with a message
    before anything touches a device."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_bench_reads_the_committed_traffic_figures():
    """The bench lines' measured HBM traffic comes from the committed rocprofv3
    summaries: both figures parse, and a missing file or key gives none."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_for_test", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    b, src = bench.committed_figure(os.path.join("profiles", "queued_batch_rmat22_traffic.md"),
                                    "traffic_bytes_per_batch")
    assert b is not None and b > 0 and src.endswith("queued_batch_rmat22_traffic.md")
    c, src = bench.committed_figure(os.path.join("profiles", "compaction_rmat22_traffic.md"),
                                    "traffic_bytes_per_compaction")
    assert c is not None and c > b and src.endswith("compaction_rmat22_traffic.md")
    assert bench.committed_figure(os.path.join("profiles", "no_such_summary.md"), "traffic") == (None, None)
    assert bench.committed_figure(os.path.join("profiles", "compaction_rmat22_traffic.md"), "no_such_key") == \
        (None, None)


def test_binaries_usage_without_gpu(engine_mod):
    for b in ("misfit", "misfit_async"):
        exe = os.path.join(engine_mod.BIN_DIR, b)
        assert os.path.exists(exe), exe
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 1
        assert "Usage:" in r.stderr and "<num_batches>" in r.stderr


# ------------------------------------------------------------ .mtx -> .egr and the validator
def test_converter_reproduces_reference_sample(engine_mod, tmp_path):
    """sample.egr in the reference's dataset is the converter's output for
    sample.mtx (tools/efficient_converter.py): ours must give the same bytes."""
    out = tmp_path / "s.egr"
    n, m = engine_mod.mtx_to_egr(os.path.join(SAMPLE, "sample.mtx"), str(out))
    assert (n, m) == (2000, 16000)
    assert out.read_bytes() == open(os.path.join(SAMPLE, "sample.egr"), "rb").read()


def test_converter_semantics(engine_mod, tmp_path):
    """Dimension line = max(rows, cols); rows in file order; u > n dropped;
    short lines and later comments skipped; no symmetrisation or dedup."""
    src = tmp_path / "g.mtx"
    src.write_text("%%MatrixMarket matrix coordinate pattern general\n% comment\n"
                   "3 4\n"            # no entry count: n = 4
                   "2 3\n1 4\n2 1\n"
                   "% inner comment\n"
                   "9 1\n"            # u > n: dropped
                   "2 3\n"            # duplicate kept
                   "4 4\n"            # self-loop kept
                   "7\n")             # one field: skipped
    out = tmp_path / "g.egr"
    n, m = engine_mod.mtx_to_egr(str(src), str(out))
    assert (n, m) == (4, 5)
    ni, nl, wide = engine_mod.read_egr(str(out))
    assert wide
    assert ni.tolist() == [0, 1, 4, 4, 5]
    assert nl.tolist() == [3, 2, 0, 2, 3]
    rep = engine_mod.egr_check(str(out))
    assert rep["duplicates"] == 1 and rep["self_loops"] == 1 and rep["unsorted_rows"] == 1
    assert rep["asymmetric"] == -1 and rep["bad_offsets"] == 0 and rep["out_of_range"] == 0
    bad = tmp_path / "z.mtx"
    bad.write_text("2 2\n0 1\n")
    with pytest.raises(engine_mod.MisfitError):
        engine_mod.mtx_to_egr(str(bad), str(tmp_path / "z.egr"))


def test_validator_flags_asymmetry_and_range(engine_mod, tmp_path):
    ni = np.array([0, 2, 3, 3], np.int64)
    nl = np.array([1, 2, 0], np.int32)       # 0-1 both ways, 0-2 one way
    p = tmp_path / "a.egr"
    engine_mod.write_egr(str(p), ni, nl)
    rep = engine_mod.egr_check(str(p))
    assert rep["asymmetric"] == 1 and rep["unsorted_rows"] == 0
    nl2 = np.array([1, 5, 0], np.int32)      # 5 is outside [0, 3)
    engine_mod.write_egr(str(p), ni, nl2)
    rep = engine_mod.egr_check(str(p))
    assert rep["out_of_range"] == 1 and rep["asymmetric"] == -1


def test_egr_tool_binary(engine_mod, tmp_path):
    exe = os.path.join(ROOT, "misfit_amd", "bin", "misfit_egr")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    out = tmp_path / "s.egr"
    r = subprocess.run([exe, "convert", os.path.join(SAMPLE, "sample.mtx"), str(out)], capture_output=True, text=True)
    assert r.returncode == 0 and "2000 nodes, 16000 edges" in r.stdout
    r = subprocess.run([exe, "check", str(out)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")


def test_egr_tool_mis_and_batches(engine_mod, tmp_path):
    """misfit_egr mis gives the sample's own initial MIS file
    (tools/compute_initial_mis.py made sample_mis.txt from sample.egr);
    misfit_egr batches writes <k>.mtx files of distinct existing edges,
    1-indexed with u < v, which the drop-in binaries' batch parser reads."""
    exe = os.path.join(ROOT, "misfit_amd", "bin", "misfit_egr")
    out = tmp_path / "mis.txt"
    r = subprocess.run([exe, "mis", os.path.join(SAMPLE, "sample.egr"), str(out)], capture_output=True, text=True)
    assert r.returncode == 0 and "greedy MIS: 533 members" in r.stdout
    want = sorted(int(x) for x in open(os.path.join(SAMPLE, "sample_mis.txt")).read().split())
    assert [int(x) for x in out.read_text().split()] == want
    ni, nl, _ = engine_mod.read_egr(os.path.join(SAMPLE, "sample.egr"))
    adj = {(int(a), int(b)) for a in range(len(ni) - 1) for b in nl[ni[a]:ni[a + 1]]}
    for sampler in ("uniform", "low-degree", "preferential"):
        d = tmp_path / sampler
        r = subprocess.run([exe, "batches", os.path.join(SAMPLE, "sample.egr"), str(d), "100", "3", sampler, "5"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        seen = set()
        for k in (1, 2, 3):
            lines = (d / ("%d.mtx" % k)).read_text().split("\n")[:-1]
            assert len(lines) == 100
            for ln in lines:
                a, b = (int(x) for x in ln.split())
                assert 1 <= a < b <= 2000 and (a - 1, b - 1) in adj and (a, b) not in seen
                seen.add((a, b))
    r = subprocess.run([exe, "batches", os.path.join(SAMPLE, "sample.egr"), str(tmp_path / "x"), "9000", "1"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "distinct edges" in r.stderr       # 8,000 undirected edges only
    r = subprocess.run([exe, "batches", os.path.join(SAMPLE, "sample.egr"), str(tmp_path / "x"), "10", "1", "bogus"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_reference_samplers_bias(engine_mod):
    """The three existing-edge samplers: distinct (lo, hi) existing edges;
    drawing the vertex by 1/degree lowers the mean degree of the drawn
    vertex's side well below the entry-uniform and degree-weighted
    samplers, which draw the same distribution (a uniform CSR entry)."""
    ni, nl = engine_mod.gen_rmat(12, 8, 3)
    deg = np.diff(ni)
    adj = set(zip(np.repeat(np.arange(len(deg)), deg).tolist(), nl.tolist()))
    means = {}
    for kind in (engine_mod.EDGE_UNIFORM, engine_mod.EDGE_LOW_DEGREE, engine_mod.EDGE_PREFERENTIAL):
        u, v = engine_mod.gen_updates(kind, ni, nl, 3000, 9)
        assert np.all(u < v)
        assert len(set(zip(u.tolist(), v.tolist()))) == 3000
        assert all((a, b) in adj for a, b in zip(u.tolist(), v.tolist()))
        means[kind] = float(np.minimum(deg[u], deg[v]).mean())
    assert means[engine_mod.EDGE_LOW_DEGREE] < 0.5 * means[engine_mod.EDGE_UNIFORM]
    assert 0.8 < means[engine_mod.EDGE_PREFERENTIAL] / means[engine_mod.EDGE_UNIFORM] < 1.25
    with pytest.raises(engine_mod.MisfitError):
        engine_mod.gen_updates(engine_mod.EDGE_UNIFORM, ni, nl, len(nl) + 1, 1)


def test_run_script_arguments():
    """misfit_amd/run.sh keeps the reference's flags (run.sh:17-127)."""
    sh = os.path.join(ROOT, "misfit_amd", "run.sh")
    r = subprocess.run(["bash", sh, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--variant=misfit|misfit_async" in r.stderr
    r = subprocess.run(["bash", sh], capture_output=True, text=True)
    assert r.returncode == 1 and "--variant" in r.stderr
    r = subprocess.run(["bash", sh, "--variant=other", "--sample"], capture_output=True, text=True)
    assert r.returncode == 1 and "must be 'misfit' or 'misfit_async'" in r.stderr
    r = subprocess.run(["bash", sh, "--variant=misfit", "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unrecognized argument" in r.stderr
    r = subprocess.run(["bash", sh, "--variant=misfit", "--graph=/nonexistent.egr", "--mis=/x", "--batch=/y"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "graph not found" in r.stderr
    r = subprocess.run(["bash", sh, "--variant", "misfit", "--sample", "--num-batches=x"], capture_output=True, text=True)
    assert r.returncode == 1 and "--num-batches" in r.stderr
    r = subprocess.run(["bash", sh, "--variant=misfit", "--sample", "--gpu"], capture_output=True, text=True)
    assert r.returncode == 1 and "--gpu needs a value" in r.stderr


def _build_shim_sample(out_dir):
    """Compiles tests/shim/shim_sample.cpp (the reference-shaped engine class
    over the C ABI) against the built library."""
    lib_dir = os.path.join(ROOT, "misfit_amd")
    exe = os.path.join(str(out_dir), "shim_sample")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-o", exe,
                        os.path.join(ROOT, "tests", "shim", "shim_sample.cpp"), "-L" + lib_dir, "-lmisfit_amd",
                        "-Wl,-rpath," + lib_dir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_engine_class_shim_compiles(engine_mod, tmp_path):
    """misfit_amd/shim/gpu_vertex_clustering_graph.h (the reference's
    GPUVertexClusteringGraph interface, src/misfit/misfit.cu:360-693, over the
    C ABI) compiles and links as C++17 with a caller's own edge and timing
    records; the GPU test runs it on the sample."""
    assert os.path.exists(_build_shim_sample(tmp_path))
