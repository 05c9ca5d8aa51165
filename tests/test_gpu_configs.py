"""The benchmark configurations beyond the headline one, on the GPU.

* Configuration 3 shape (2-D random geometric graph, int64-sentinel .egr,
  update endpoints drawn proportionally to degree+1): bit-exact against the
  oracle at a size the oracle replays in seconds, and at the full size
  (2^24 vertices, ~537M entries, 50,000-update batches) through the
  size-independent properties: independence, maximality and count
  consistency after every batch, and the same MIS from both drivers.
* Configuration 5 shape (several graphs of mixed types held on the device at
  once, low-degree update streams): every graph's MIS after every batch is
  the same when its batches run alone and when they are interleaved with the
  other graphs' batches on the GPU, and equals the oracle's.
"""

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _batches(u, v, ins, size):
    return [(u[i:i + size], v[i:i + size], ins[i:i + size]) for i in range(0, len(u), size)]


def test_rgg_wide_form_parity(oracle, engine_mod, tmp_path):
    """RGG n = 2^15, degree 32, seed 11, through the int64 .egr form;
    4 batches of 4,000 updates with endpoints ∝ (degree + 1)."""
    ni, nl = engine_mod.gen_rgg(1 << 15, 32.0, 11)
    p = str(tmp_path / "rgg.egr")
    engine_mod.write_egr(p, ni, nl)
    ni2, nl2, wide = engine_mod.read_egr(p)
    assert wide and np.array_equal(ni2, ni) and np.array_equal(nl2, nl)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.PREFERENTIAL, ni, nl, 16000, 11)
    keep = u != v
    u, v = u[keep], v[keep]
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    batches = _batches(u, v, ins, 4000)
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    with engine_mod.Engine(0) as eng:
        eng.load_egr(p)
        for mode in (engine_mod.MODE_QUEUED, engine_mod.MODE_PLAIN):
            got, gst, checks = H.engine_replay(engine_mod, eng, ni2, nl2, mis0, batches, mode=mode)
            for b in range(len(batches)):
                assert np.array_equal(want[b], got[b]), "mode %d batch %d" % (mode, b)
                assert checks[b] == (0, 0, 0)
            assert H.counter_mismatch(wst, gst) == "", "mode %d" % mode


def test_low_degree_stream_parity(oracle, engine_mod):
    """R-MAT scale 15 with the low-degree-quartile update stream."""
    ni, nl = engine_mod.gen_rmat(15, 16, 5)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.LOW_DEGREE, ni, nl, 12000, 5)
    keep = u != v
    u, v = u[keep], v[keep]
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    batches = _batches(u, v, ins, 3000)
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    with engine_mod.Engine(0) as eng:
        got, gst, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches)
    for b in range(len(batches)):
        assert np.array_equal(want[b], got[b]), "batch %d" % b
        assert checks[b] == (0, 0, 0)
    assert H.counter_mismatch(wst, gst) == ""


def test_rgg_full_size_properties(engine_mod):
    """Configuration 3 at full size: n = 2^24, average degree 32, seed 11,
    int64 nindex; 3 batches of 50,000 updates ∝ (degree + 1).  Checked after
    every batch; both drivers must end with the same MIS."""
    print("generating RGG 2^24", flush=True)
    ni, nl = engine_mod.gen_rgg(1 << 24, 32.0, 11)
    assert len(nl) > 500_000_000
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.PREFERENTIAL, ni, nl, 150000, 11)
    keep = u != v
    u, v = u[keep], v[keep]
    finals = []
    with engine_mod.Engine(0) as eng:
        for mode in (engine_mod.MODE_QUEUED, engine_mod.MODE_PLAIN):
            print("replaying, mode", mode, flush=True)
            eng.load_csr(ni, nl)
            ins = eng.flags(u, v)
            ups = engine_mod.make_updates(u, v, ins)
            eng.set_mis(mis0)
            eng.set_mode(mode)
            for b in range(0, len(ups), 50000):
                st = eng.process_batch(ups[b:b + 50000])
                assert st["updates"] == min(50000, len(ups) - b)
                assert eng.check() == (0, 0, 0), "mode %d batch %d" % (mode, b // 50000)
            finals.append(eng.get_mis())
    assert np.array_equal(finals[0], finals[1])


def test_graphs_alone_and_in_the_mix(oracle, engine_mod):
    """Six graphs of mixed types on the device at once.  Each graph's batches
    give the same MIS alone and interleaved with the others (all contexts
    queued on their own streams, submitted round-robin before any wait), and
    the oracle's."""
    specs = [("er", 1 << 14, 1), ("rmat", 14, 2), ("rgg", 1 << 14, 3), ("er", 1 << 15, 4),
             ("rmat", 15, 5), ("rgg", 1 << 15, 6)]
    graphs = []
    for kind, size, seed in specs:
        if kind == "er":
            ni, nl = engine_mod.gen_er(size, 8 * size, seed)
        elif kind == "rmat":
            ni, nl = engine_mod.gen_rmat(size, 16, seed)
        else:
            ni, nl = engine_mod.gen_rgg(size, 32.0, seed)
        mis0 = engine_mod.greedy_mis(ni, nl)
        u, v = engine_mod.gen_updates(engine_mod.LOW_DEGREE, ni, nl, 6000, seed)
        keep = u != v
        u, v = u[keep], v[keep]
        ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
        batches = _batches(u, v, ins, 2000)
        want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
        graphs.append((ni, nl, mis0, batches, want))
    # alone
    alone = []
    with engine_mod.Engine(0) as eng:
        for (ni, nl, mis0, batches, want) in graphs:
            got, _, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches)
            alone.append(got)
            for b in range(len(batches)):
                assert np.array_equal(got[b], want[b])
                assert checks[b] == (0, 0, 0)
    # in the mix: one context per graph, batches interleaved across graphs
    engines = [engine_mod.Engine(0) for _ in graphs]
    try:
        for eng, (ni, nl, mis0, batches, want) in zip(engines, graphs):
            eng.load_csr(ni, nl)
            eng.set_mis(mis0)
            eng.set_mode(engine_mod.MODE_QUEUED)
        nb = max(len(g[3]) for g in graphs)
        for b in range(nb):
            tickets = []
            for gi, (eng, g) in enumerate(zip(engines, graphs)):
                if b < len(g[3]):
                    u, v, ins = g[3][b]
                    tickets.append((gi, eng, eng.submit(engine_mod.make_updates(u, v, ins))))
            for gi, eng, t in tickets:
                eng.wait(t)
                assert np.array_equal(eng.get_mis(), alone[gi][b]), "graph %d batch %d" % (gi, b)
        for eng in engines:
            assert eng.check() == (0, 0, 0)
    finally:
        for eng in engines:
            eng.close()
