"""The serving layer (misfit_server_*) on the GPU.

Requests for several graphs are submitted interleaved and served
concurrently; graphs are compacted when their overlay passes a bound.  Every
graph must end with the MIS the oracle computes when it replays the same
batches with compactions at the same points (each request reports how many
compactions of its graph ran before it), for every in-flight depth.  Graphs
and batches that do not fit in free device memory minus the reserve are
refused without changing anything.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _workload(engine_mod, oracle, kind, size, seed, nb=5, bs=3000):
    if kind == "er":
        ni, nl = engine_mod.gen_er(size, 8 * size, seed)
    elif kind == "rmat":
        ni, nl = engine_mod.gen_rmat(size, 16, seed)
    else:
        ni, nl = engine_mod.gen_rgg(size, 32.0, seed)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, nb * bs, seed)
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    batches = [(u[i:i + bs], v[i:i + bs], ins[i:i + bs]) for i in range(0, nb * bs, bs)]
    return ni, nl, mis0, batches


def _oracle_with_compactions(oracle, ni, nl, mis0, batches, before, stats=None):
    """Replay with a compaction wherever `before` (compactions before each
    batch) steps up; each batch's rule counters go to `stats` when given."""
    g = oracle.Graph.from_csr(ni, nl)
    mis = np.array(mis0, np.uint8)
    done = 0
    for (u, v, ins), cb in zip(batches, before):
        while done < cb:
            cni, cnl = g.compact()
            g = oracle.Graph.from_csr(cni, cnl)
            done += 1
        st = g.process_batch(mis, u, v, ins, completion=True, literal=False)
        if stats is not None:
            stats.append(st)
    return mis


@pytest.mark.parametrize("inflight,keep", [(1, "0"), (3, "0"), (1, "1")])
def test_served_graphs_match_oracle_with_compactions(oracle, engine_mod, inflight, keep, monkeypatch):
    """Three graphs served with compactions; keep "1": every graph keeps the
    buffers its compactions release (MISFIT_SERVE_COMPACT_KEEP, read when
    the server is created) and takes its next compaction's outputs from them."""
    monkeypatch.setenv("MISFIT_SERVE_COMPACT_KEEP", keep)
    work = [_workload(engine_mod, oracle, k, s, seed)
            for (k, s, seed) in [("er", 1 << 14, 31), ("rmat", 14, 32), ("rgg", 1 << 14, 33)]]
    with engine_mod.Server(0, compact_fraction=0.02, max_inflight=inflight) as srv:
        gids = [srv.add_graph(ni, nl, mis0) for (ni, nl, mis0, _) in work]
        reqs = {g: [] for g in gids}
        for b in range(5):
            for gid, (_, _, _, batches) in zip(gids, work):
                u, v, ins = batches[b]
                reqs[gid].append(srv.submit(gid, engine_mod.make_updates(u, v, ins)))
        for gid, (ni, nl, mis0, batches) in zip(gids, work):
            stats = [srv.wait(r) for r in reqs[gid]]
            before = [s["compactions_before"] for s in stats]
            assert before == sorted(before)
            for s in stats:
                assert s["graph"] == gid and s["updates"] == 3000
                assert 0 <= s["queue_ms"] <= s["latency_ms"] and s["device_ms"] > 0
            ost = []
            want = _oracle_with_compactions(oracle, ni, nl, mis0, batches, before, ost)
            got = srv.get_mis(gid, len(ni) - 1)
            assert np.array_equal(got, want), "graph %d" % gid
            assert srv.check(gid) == (0, 0, 0)
            assert H.counter_mismatch(ost, [s["batch"] for s in stats]) == "", "graph %d" % gid
        # the bound is low enough that compactions did run
        assert any(srv.wait(srv.submit(g, engine_mod.make_updates(np.zeros(0, np.int32), np.zeros(0, np.int32),
                                                                    np.zeros(0, np.uint8))))["compactions_before"] > 0
                   for g in gids)


@pytest.mark.parametrize("inflight,workers", [(1, 0), (1, 1), (2, 2), (3, 0)])
def test_served_dense_conflict_graph(oracle, engine_mod, inflight, workers):
    """A graph whose first batches need far more rounds than first queued
    (8,192-vertex R-MAT, 8,000-update batches of low-degree endpoints),
    served beside a larger graph: its batches wait while one settles its
    round count, and both graphs end with the oracle's MIS."""
    ni, nl = engine_mod.gen_rmat(13, 16, 21)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.LOW_DEGREE, ni, nl, 4 * 8000, 21)
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    dense = (ni, nl, mis0, [(u[i:i + 8000], v[i:i + 8000], ins[i:i + 8000]) for i in range(0, 32000, 8000)])
    other = _workload(engine_mod, oracle, "er", 1 << 15, 44, nb=4)
    with engine_mod.Server(0, compact_fraction=0.0, max_inflight=inflight, workers=workers) as srv:
        gids = [srv.add_graph(w[0], w[1], w[2]) for w in (dense, other)]
        reqs = []
        for b in range(4):
            for gid, w in zip(gids, (dense, other)):
                reqs.append((gid, srv.submit(gid, engine_mod.make_updates(*w[3][b]))))
        for gid, r in reqs:
            assert srv.wait(r)["graph"] == gid
        for gid, w in zip(gids, (dense, other)):
            want, _, _ = H.oracle_replay(oracle, w[0], w[1], w[2], w[3], completion=True, literal=False)
            assert np.array_equal(srv.get_mis(gid, len(w[0]) - 1), want[-1]), "graph %d" % gid
            assert srv.check(gid) == (0, 0, 0)


def test_admission_by_free_memory(oracle, engine_mod):
    ni, nl, mis0, batches = _workload(engine_mod, oracle, "er", 1 << 14, 41, nb=1)
    with engine_mod.Engine(0) as probe:
        _, total, free = probe.device_info()
    # a reserve larger than the device: nothing fits
    with engine_mod.Server(0, hbm_reserve=total + (1 << 30)) as srv:
        with pytest.raises(engine_mod.MisfitError) as ei:
            srv.add_graph(ni, nl, mis0)
        assert ei.value.code == -4
    # the graph fits; a batch whose buffers would cut into the reserve is
    # refused and leaves the graph as it was, a small one is served
    with engine_mod.Server(0) as srv:
        gid = srv.add_graph(ni, nl, mis0)
        with engine_mod.Engine(0) as probe:
            _, _, free_now = probe.device_info()
    with engine_mod.Server(0, hbm_reserve=max(free_now - (300 << 20), 0)) as srv:
        gid = srv.add_graph(ni, nl, mis0)
        big = np.zeros(400_000, engine_mod.UPDATE_DTYPE)
        big["u"] = 0
        big["v"] = 1
        with pytest.raises(engine_mod.MisfitError) as ei:
            srv.wait(srv.submit(gid, big))
        assert ei.value.code == -4
        assert np.array_equal(srv.get_mis(gid, len(ni) - 1), mis0)
        u, v, ins = batches[0]
        st = srv.wait(srv.submit(gid, engine_mod.make_updates(u, v, ins)))
        assert st["updates"] == len(u)
        want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
        assert np.array_equal(srv.get_mis(gid, len(ni) - 1), want[0])


def test_compaction_without_headroom_fails_the_next_request(oracle, engine_mod):
    """A compaction whose headroom does not fit in free memory minus the
    reserve does not run: the graph stays as it was, the next request fails
    with MISFIT_E_NOMEM without being applied, and once the reserve allows it
    the compaction runs after the overlay has grown further; the MIS equals
    the oracle's with compactions at the reported points."""
    ni, nl, mis0, batches = _workload(engine_mod, oracle, "er", 1 << 14, 61, nb=3)
    with engine_mod.Server(0, compact_fraction=0.005, max_inflight=1) as srv:
        gid = srv.add_graph(ni, nl, mis0)
        with engine_mod.Engine(0) as probe:
            _, _, free_now = probe.device_info()
        # room for a 3,000-update batch (3 MB), none for the compaction
        # (about 20 MiB for this graph)
        srv.set_reserve(max(free_now - (12 << 20), 0))
        st1 = srv.wait(srv.submit(gid, engine_mod.make_updates(*batches[0])))
        assert st1["compactions_before"] == 0
        with pytest.raises(engine_mod.MisfitError) as ei:
            srv.wait(srv.submit(gid, engine_mod.make_updates(*batches[1])))
        assert ei.value.code == -4 and "compaction" in str(ei.value)
        want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches[:1], completion=True, literal=False)
        assert np.array_equal(srv.get_mis(gid, len(ni) - 1), want[0])
        srv.set_reserve(0)
        st2 = srv.wait(srv.submit(gid, engine_mod.make_updates(*batches[1])))
        st3 = srv.wait(srv.submit(gid, engine_mod.make_updates(*batches[2])))
        before = [st1["compactions_before"], st2["compactions_before"], st3["compactions_before"]]
        assert before == [0, 0, 1]
        want = _oracle_with_compactions(oracle, ni, nl, mis0, batches, before)
        assert np.array_equal(srv.get_mis(gid, len(ni) - 1), want)
        assert srv.check(gid) == (0, 0, 0)


def test_concurrent_clients_and_graph_churn(oracle, engine_mod):
    """Client threads submit to their own graphs at the same time while a
    graph is removed and another added; every graph ends with the oracle's
    MIS (one batch per graph in flight, four dispatcher threads)."""
    import threading
    work = [_workload(engine_mod, oracle, k, s, seed, nb=4, bs=2000)
            for (k, s, seed) in [("er", 1 << 13, 51), ("rmat", 13, 52), ("rgg", 1 << 13, 53), ("er", 1 << 12, 54)]]
    with engine_mod.Server(0, compact_fraction=0.0, max_inflight=1, workers=4) as srv:
        gids = [srv.add_graph(ni, nl, mis0) for (ni, nl, mis0, _) in work[:3]]
        errors = []

        def client(gid, batches):
            try:
                reqs = [srv.submit(gid, engine_mod.make_updates(*b)) for b in batches]
                for r in reqs:
                    srv.wait(r)
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)

        threads = [threading.Thread(target=client, args=(g, w[3])) for g, w in zip(gids, work[:3])]
        for t in threads:
            t.start()
        # churn while the clients run: a fourth graph comes and goes, then comes back
        g4 = srv.add_graph(*work[3][:3])
        srv.remove_graph(g4)
        g4 = srv.add_graph(*work[3][:3])
        client(g4, work[3][3])
        for t in threads:
            t.join()
        assert not errors, errors
        for gid, (ni, nl, mis0, batches) in zip(gids + [g4], work):
            want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
            assert np.array_equal(srv.get_mis(gid, len(ni) - 1), want[-1]), "graph %d" % gid
            assert srv.check(gid) == (0, 0, 0)


def test_tuner_on_the_device(engine_mod, tmp_path, monkeypatch):
    """The autotuner times the queued driver under each knob setting on a
    small R-MAT stream: every trial ends with the MIS of the defaults (the
    tuner checks it), which is also the plain driver's, and the tuned
    settings load back into the environment."""
    from misfit_amd import tune as T
    for k in T.SPACE:
        # set, then delete: the knobs the tuner's apply() sets are undone after the test
        monkeypatch.setenv(k, "0")
        monkeypatch.delenv(k)
    ni, nl = engine_mod.gen_rmat(14, 16, 5)
    mis0 = engine_mod.greedy_mis(ni, nl)
    B, W, K = 2000, 2, 6
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, (W + K) * B, 5)
    with engine_mod.Engine(0) as eng:
        eng.load_csr(ni, nl)
        ups = engine_mod.make_updates(u, v, eng.flags(u, v))
        eng.set_mis(mis0)
        eng.set_mode(engine_mod.MODE_PLAIN)
        for b in range(W + K):
            eng.process_batch(ups[b * B:(b + 1) * B])
        want = eng.get_mis()
    space = {"MISFIT_ROUND_MARGIN": ["", "1", "4"], "MISFIT_WINDOW_FACTOR": ["", "8"],
             "MISFIT_ROUND_WG_PER_CU": ["", "2"], "MISFIT_ROUND_THREADS": ["", "128", "256"],
             "MISFIT_PU_DENSE_FRACTION": ["", "0"]}
    seen = []

    def fn(env):
        r, mis = T.measure(engine_mod, ni, nl, mis0, ups, B, W, K, env)
        seen.append(mis)
        return r, mis
    res = T.tune(fn, space=space, repeats=1)
    assert len(res["trials"]) == 1 + 2 + 1 + 1 + 2 + 1
    assert all(np.array_equal(m, want) for m in seen)
    assert res["updates_per_s"] >= res["default_updates_per_s"] > 0
    p = str(tmp_path / "tune.json")
    T.save(res, p)
    assert T.apply(p) == res["env"]


def _mixed_graphs(engine_mod, oracle, count, nb, seed0):
    """`count` small graphs of the three types, with low-degree update
    batches of uneven sizes (dense contention on the smallest graphs)."""
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(count):
        kind = ("er", "rmat", "rgg")[i % 3]
        size = int(rng.integers(12, 15))
        seed = seed0 * 100 + i
        if kind == "er":
            ni, nl = engine_mod.gen_er(1 << size, 8 << size, seed)
        elif kind == "rmat":
            ni, nl = engine_mod.gen_rmat(size, 16, seed)
        else:
            ni, nl = engine_mod.gen_rgg(1 << size, 32.0, seed)
        mis0 = engine_mod.greedy_mis(ni, nl)
        sizes = [int(x) for x in rng.integers(200, 4000, nb)]
        u, v = engine_mod.gen_updates(engine_mod.LOW_DEGREE, ni, nl, sum(sizes), seed)
        ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
        offs = np.concatenate([[0], np.cumsum(sizes)])
        batches = [(u[a:b], v[a:b], ins[a:b]) for a, b in zip(offs[:-1], offs[1:])]
        out.append((ni, nl, mis0, batches))
    return out


def _serve_all(engine_mod, graphs, **opts):
    """Every request submitted at once, interleaved across graphs; per graph
    the MIS after all its batches, the request stats and the final check."""
    with engine_mod.Server(0, **opts) as srv:
        gids = [srv.add_graph(ni, nl, mis0) for (ni, nl, mis0, _) in graphs]
        nb = max(len(g[3]) for g in graphs)
        reqs = [(gi, srv.submit(gids[gi], engine_mod.make_updates(*graphs[gi][3][b])))
                for b in range(nb) for gi in range(len(graphs)) if b < len(graphs[gi][3])]
        stats = {gi: [] for gi in range(len(graphs))}
        for gi, r in reqs:
            stats[gi].append(srv.wait(r))
        mis = [srv.get_mis(gids[gi], len(g[0]) - 1) for gi, g in enumerate(graphs)]
        checks = [srv.check(g) for g in gids]
    return mis, stats, checks


def test_serving_kernel_graphs_alone_and_in_the_mix(oracle, engine_mod):
    """Sixteen graphs whose batches all run in the serving kernel (each
    batch whole in one workgroup, workgroups taking batches of any graph
    from one queue): each graph's MIS is the same served alone and in the
    mix of all sixteen, and the oracle's."""
    graphs = _mixed_graphs(engine_mod, oracle, 16, 3, 7)
    reps = [H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
            for (ni, nl, mis0, batches) in graphs]
    want = [r[0][-1] for r in reps]
    mix, stats, checks = _serve_all(engine_mod, graphs, compact_fraction=0.0, max_inflight=1, workers=4,
                                    persistent=2)
    for gi in range(len(graphs)):
        assert np.array_equal(mix[gi], want[gi]), "graph %d in the mix" % gi
        assert checks[gi] == (0, 0, 0)
        assert H.counter_mismatch(reps[gi][1], [s["batch"] for s in stats[gi]]) == "", "graph %d" % gi
        assert all(s["device_ms"] > 0 and s["batch"]["rounds"] >= 0 for s in stats[gi])
    for gi in (0, 5, 10, 15):  # alone: a server holding only this graph
        alone, _, checks1 = _serve_all(engine_mod, graphs[gi:gi + 1], compact_fraction=0.0, max_inflight=1,
                                       persistent=2)
        assert np.array_equal(alone[0], mix[gi]), "graph %d alone" % gi
        assert checks1[0] == (0, 0, 0)


def test_serving_kernel_beside_streams_with_compactions(oracle, engine_mod):
    """Batches split between the serving kernel (small or contended ones)
    and the graphs' own streams (the rest), with compactions between them:
    every graph ends with the MIS of the oracle compacting at the reported
    points."""
    graphs = _mixed_graphs(engine_mod, oracle, 9, 4, 11)
    # two larger, uncontended graphs whose big batches take the streams
    for kind, size, seed in (("er", 1 << 15, 91), ("rmat", 15, 92)):
        graphs.append(_workload(engine_mod, oracle, kind, size, seed, nb=4, bs=6000))
    mis, stats, checks = _serve_all(engine_mod, graphs, compact_fraction=0.01, max_inflight=1, workers=4,
                                    persistent=1)
    compacted = 0
    for gi, (ni, nl, mis0, batches) in enumerate(graphs):
        before = [s["compactions_before"] for s in stats[gi]]
        compacted += before[-1]
        ost = []
        want = _oracle_with_compactions(oracle, ni, nl, mis0, batches, before, ost)
        assert np.array_equal(mis[gi], want), "graph %d" % gi
        assert checks[gi] == (0, 0, 0)
        assert H.counter_mismatch(ost, [s["batch"] for s in stats[gi]]) == "", "graph %d" % gi
    assert compacted > 0
