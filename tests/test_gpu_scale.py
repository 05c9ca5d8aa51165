"""The benchmark's and the serving configuration's workloads at full size on
the GPU, and the serving configuration scaled down against the oracle.

* R-MAT scale 22 (128,303,020 entries, int32 offsets in the .egr form),
  10,000-update batches: the plain driver (host-driven rounds) checked after
  every batch, the queued driver with two batches in flight, equal MIS.
* R-MAT scale 26 (2.1 B entries, int64 offsets on the device) through the
  serving layer: batch sizes log-uniform in [100, 10^6], half deletions of
  existing edges and half insertions of non-edges, compaction when the
  overlay passes 5% of the base; the full check after every batch, at least
  one compaction, and the final MIS equal to the plain driver's replaying the
  same batches with the compactions at the same points.
* R-MAT scale 16 through the serving layer with batches up to 10^5 updates
  and compactions, bit-exact with the oracle compacting at the same points.
"""
import time

import numpy as np
import pytest

from tests import helpers as H
from tests.test_gpu_serve import _oracle_with_compactions

pytestmark = pytest.mark.gpu


def _log_uniform_sizes(rng, lo, hi, total):
    sizes = []
    while sum(sizes) < total:
        sizes.append(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    return sizes


def _say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(600)
def test_rmat22_stream_plain_checked_queued_equal(engine_mod):
    nb, bs = 12, 10000
    t = time.time()
    ni, nl = engine_mod.gen_rmat(22, 16, 7)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, nb * bs, 7)
    _say("R-MAT 22: %d entries, %d updates generated in %.1f s" % (len(nl), nb * bs, time.time() - t))
    assert len(nl) == 128_303_020
    finals, counters = {}, {mode: [] for mode in (engine_mod.MODE_PLAIN, engine_mod.MODE_QUEUED)}
    for mode in (engine_mod.MODE_PLAIN, engine_mod.MODE_QUEUED):
        with engine_mod.Engine(0) as eng:
            eng.load_csr(ni, nl)
            ups = engine_mod.make_updates(u, v, eng.flags(u, v))
            eng.set_mis(mis0)
            eng.set_mode(mode)
            if mode == engine_mod.MODE_PLAIN:
                for b in range(nb):
                    st = eng.process_batch(ups[b * bs:(b + 1) * bs])
                    assert st["updates"] == bs
                    counters[mode].append(st)
                    assert eng.check() == (0, 0, 0), "plain batch %d" % b
            else:
                tickets = []
                for b in range(nb):
                    tickets.append(eng.submit(ups[b * bs:(b + 1) * bs]))
                    if len(tickets) == 2:
                        counters[mode].append(eng.wait(tickets.pop(0)))
                for tk in tickets:
                    counters[mode].append(eng.wait(tk))
                assert eng.check() == (0, 0, 0)
            finals[mode] = eng.get_mis()
        _say("mode %d done" % mode)
    assert np.array_equal(finals[engine_mod.MODE_PLAIN], finals[engine_mod.MODE_QUEUED])
    # the rule's counters per batch agree between the drivers too
    assert H.counter_mismatch(counters[engine_mod.MODE_PLAIN], counters[engine_mod.MODE_QUEUED]) == ""
    assert sum(s["evictions"] for s in counters[engine_mod.MODE_PLAIN]) > 0


@pytest.mark.timeout(900)
def test_rmat16_serving_large_batches_with_compactions_match_oracle(oracle, engine_mod):
    ni, nl = engine_mod.gen_rmat(16, 16, 3)
    mis0 = engine_mod.greedy_mis(ni, nl)
    sizes = [100, 2000, 30000, 100000, 500, 60000, 7000]
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, sum(sizes), 3)
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    batches = [(u[a:b], v[a:b], ins[a:b]) for a, b in zip(offs[:-1], offs[1:])]
    with engine_mod.Server(0, compact_fraction=0.02, max_inflight=2) as srv:
        gid = srv.add_graph(ni, nl, mis0)
        reqs = [srv.submit(gid, engine_mod.make_updates(*b)) for b in batches]
        stats = [srv.wait(r) for r in reqs]
        got = srv.get_mis(gid, len(ni) - 1)
        assert srv.check(gid) == (0, 0, 0)
    before = [s["compactions_before"] for s in stats]
    _say("compactions before each batch: %s" % before)
    assert before == sorted(before) and before[-1] >= 1
    ost = []
    want = _oracle_with_compactions(oracle, ni, nl, mis0, batches, before, ost)
    assert np.array_equal(got, want)
    assert H.counter_mismatch(ost, [s["batch"] for s in stats]) == ""


@pytest.mark.timeout(1200)
def test_rmat26_serving_full_size(engine_mod):
    rng = np.random.default_rng(3)
    t = time.time()
    ni, nl = engine_mod.gen_rmat(26, 16, 3)
    n = len(ni) - 1
    _say("R-MAT 26: %d vertices, %d entries in %.1f s" % (n, len(nl), time.time() - t))
    assert n == 1 << 26 and len(nl) == 2_103_842_088  # offsets are int64 on the device
    mis0 = engine_mod.greedy_mis(ni, nl)
    # a stream long enough for the overlay to pass 5% of the base (deletions
    # of base edges take no overlay slot, insertions two or more); the
    # served run stops a few batches after the first compaction
    sizes = _log_uniform_sizes(rng, 100, 1_000_000, int(0.05 * len(nl) * 1.3))
    total = sum(sizes)
    t = time.time()
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, total, 3)
    _say("%d batches, %d updates generated in %.1f s" % (len(sizes), total, time.time() - t))
    offs = np.concatenate([[0], np.cumsum(sizes)])
    with engine_mod.Engine(0) as plain:
        plain.load_csr(ni, nl)
        ups = engine_mod.make_updates(u, v, plain.flags(u, v))
        del u, v
        plain.set_mis(mis0)
        plain.set_mode(engine_mod.MODE_PLAIN)
        with engine_mod.Server(0, compact_fraction=0.05, max_inflight=2) as srv:
            gid = srv.add_graph(ni, nl, mis0)
            before, served_stats = [], []
            t = time.time()
            for i, (a, b) in enumerate(zip(offs[:-1], offs[1:])):
                st = srv.wait(srv.submit(gid, ups[a:b]))
                assert st["updates"] == b - a
                before.append(st["compactions_before"])
                served_stats.append(st["batch"])
                assert srv.check(gid) == (0, 0, 0), "served batch %d" % i
                if i % 50 == 0:
                    _say("served batch %d / %d (%.1f s), compactions %d" % (i, len(sizes), time.time() - t,
                                                                              before[-1]))
                if before[-1] >= 1 and before.index(before[-1]) + 5 <= i:
                    break
            served = srv.get_mis(gid, n)
        nserved = len(before)
        _say("served %d batches (%d updates); compactions before the last: %d" % (nserved, offs[nserved],
                                                                                  before[-1]))
        assert before[-1] >= 1
        done = 0
        t = time.time()
        plain_stats = []
        for i in range(nserved):
            while done < before[i]:
                plain.compact()
                done += 1
            plain_stats.append(plain.process_batch(ups[offs[i]:offs[i + 1]]))
            if i % 50 == 0:
                _say("plain batch %d / %d (%.1f s)" % (i, nserved, time.time() - t))
        assert plain.check() == (0, 0, 0)
        assert np.array_equal(plain.get_mis(), served)
        # the rule's counters per batch agree between the served and the plain driver
        assert H.counter_mismatch(plain_stats, served_stats) == ""
