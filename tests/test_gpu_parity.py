"""GPU engine vs the CPU oracle, bit-exact, through the C ABI.

Every test replays the same batches through oracle/ (CPU restatement of the
reference rule) and through libmisfit_amd.so on the GPU, and compares the MIS
after every batch, the overlay's adjacency order and the compacted CSR.
"""
import numpy as np
import pytest

from tests import helpers as H
from tests.conftest import SAMPLE

pytestmark = pytest.mark.gpu

MODES = [0, 1]  # plain, queued


@pytest.fixture(scope="module")
def eng(engine_mod):
    e = engine_mod.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("completion", [False, True])
def test_sample_parity(oracle, engine_mod, eng, mode, completion):
    g, mis0, batches = H.sample_inputs(oracle, SAMPLE)
    ni, nl = g.base_csr()
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=completion, literal=True)
    got, gst, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, completion=completion, mode=mode)
    for b in range(len(batches)):
        assert np.array_equal(want[b], got[b]), "batch %d differs" % b
        assert checks[b][0] == 0 and checks[b][2] == 0
    if completion:
        assert all(c[1] == 0 for c in checks)
    else:
        # literal rule on the sample: no admissions, uncovered 4, 9, 12, 15, 19
        assert [c[1] for c in checks] == [4, 9, 12, 15, 19]
        assert int(got[-1].sum()) == 533
        assert all(s["admit_del"] == 0 and s["admit_ins"] == 0 for s in gst)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("general", [False, True])
def test_random_parity(oracle, engine_mod, eng, mode, seed, general):
    rng = np.random.default_rng(seed)
    n = 300
    ni, nl = H.random_graph(rng, n, 0.03)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 6, 80, general=general)
    for completion in (False, True):
        want, wst, og = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=completion, literal=True)
        got, gst, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, completion=completion, mode=mode)
        for b in range(len(batches)):
            assert np.array_equal(want[b], got[b]), "seed %d batch %d completion %s" % (seed, b, completion)
            assert checks[b][0] == 0 and checks[b][2] == 0
        assert H.counter_mismatch(wst, gst) == "", "seed %d completion %s" % (seed, completion)
        for v in range(0, n, 7):
            assert np.array_equal(eng.neighbors(v), og.neighbors(v)), "row %d order" % v


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", [5, 6])
def test_deletion_heavy_parity(oracle, engine_mod, eng, mode, seed):
    rng = np.random.default_rng(seed)
    ni, nl = H.random_graph(rng, 400, 0.05)
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    raw = H.existing_edge_batches(rng, ni, nl, 6, 120)
    batches = [(u, v, g.flags(u, v)) for (u, v) in raw]
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=True)
    got, gst, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, completion=True, mode=mode)
    for b in range(len(batches)):
        assert np.array_equal(want[b], got[b]), "batch %d" % b
        assert checks[b] == (0, 0, 0)
    assert H.counter_mismatch(wst, gst) == ""


def test_flags_parity(oracle, engine_mod, eng):
    rng = np.random.default_rng(9)
    ni, nl = H.random_graph(rng, 500, 0.02)
    eng.load_csr(ni, nl)
    g = oracle.Graph.from_csr(ni, nl)
    u = rng.integers(0, 500, 3000).astype(np.int32)
    v = rng.integers(0, 500, 3000).astype(np.int32)
    keep = u != v
    u, v = u[keep], v[keep]
    assert np.array_equal(eng.flags(u, v), g.flags(u, v))


@pytest.mark.parametrize("mode", MODES)
def test_compaction_parity(oracle, engine_mod, eng, mode):
    rng = np.random.default_rng(11)
    ni, nl = H.random_graph(rng, 350, 0.04)
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 4, 100, general=True)
    want, _, og = H.oracle_replay(oracle, ni, nl, mis0, batches)
    got, _, _ = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, mode=mode)
    assert np.array_equal(want[-1], got[-1])
    oni, onl = og.compact()
    m = eng.compact()
    eni, enl = eng.export_csr()
    assert m == len(onl)
    assert np.array_equal(eni, oni) and np.array_equal(enl, onl)
    # the engine keeps working on the compacted base
    more = H.random_batches(rng, oracle, oni, onl, 2, 60, general=True)
    mis = want[-1]
    w2, _, _ = H.oracle_replay(oracle, oni, onl, mis, more)
    for (u, v, ins), w in zip(more, w2):
        eng.process_batch(engine_mod.make_updates(u, v, ins))
        assert np.array_equal(eng.get_mis(), w)
    assert eng.check() == (0, 0, 0)


def test_compaction_hub_rows_across_chunks(oracle, engine_mod, eng):
    """Two hub rows of 30,000 and 20,000 neighbours, longer than the
    compaction's walk chunks (8,192 rows plus entries), with insertions into
    them and deletions from them: chunks start inside the hubs' rows and
    inside their tails' value ranges.  Compacted twice (ascending rows, then
    the merged sorted copy), each result equal to the oracle's."""
    rng = np.random.default_rng(17)
    n = 60000
    pairs = set()
    for hub, deg in ((0, 30000), (1, 20000)):
        for y in rng.choice(np.arange(2, n), deg, replace=False):
            pairs.add((hub, int(y)))
    a = rng.integers(2, n, 40000)
    b = rng.integers(2, n, 40000)
    pairs |= {(int(min(x, y)), int(max(x, y))) for x, y in zip(a, b) if x != y}
    rows = [[] for _ in range(n)]
    for x, y in pairs:
        rows[x].append(y)
        rows[y].append(x)
    ni = np.zeros(n + 1, np.int64)
    ni[1:] = np.cumsum([len(r) for r in rows])
    nl = np.concatenate([np.sort(np.array(r, np.int32)) for r in rows]).astype(np.int32)
    mis = oracle.Graph.from_csr(ni, nl).greedy_mis()
    eng.load_csr(ni, nl)
    eng.set_mis(mis)
    eng.set_rule(True)
    for k in range(2):
        g = oracle.Graph.from_csr(ni, nl)
        hub_rows = [nl[ni[h]:ni[h + 1]] for h in (0, 1)]
        batches = []
        for _ in range(2):
            # new neighbours of the hubs, deletions of hub entries, and random pairs
            hu = np.concatenate([np.zeros(600, np.int32), np.ones(400, np.int32)])
            hv = rng.integers(2, n, 1000).astype(np.int32)
            du = np.concatenate([np.zeros(500, np.int32), np.ones(300, np.int32)])
            dv = np.concatenate([rng.choice(hub_rows[0], 500), rng.choice(hub_rows[1], 300)]).astype(np.int32)
            ru = rng.integers(0, n, 2000).astype(np.int32)
            rv = rng.integers(0, n, 2000).astype(np.int32)
            u = np.concatenate([hu, du, ru])
            v = np.concatenate([hv, dv, rv])
            keep = u != v
            u, v = u[keep], v[keep]
            batches.append((u, v, g.flags(u, v)))
        want, _, og = H.oracle_replay(oracle, ni, nl, mis, batches)
        for (u, v, ins), w in zip(batches, want):
            eng.process_batch(engine_mod.make_updates(u, v, ins))
            assert np.array_equal(eng.get_mis(), w), "phase %d" % k
        oni, onl = og.compact()
        assert eng.compact() == len(onl)
        eni, enl = eng.export_csr()
        assert np.array_equal(eni, oni) and np.array_equal(enl, onl), "compaction %d" % (k + 1)
        ni, nl, mis = oni, onl, want[-1]
    assert eng.check() == (0, 0, 0)


def test_compaction_keeps_ascending_rows(oracle, engine_mod, eng):
    """Insertions that continue every row in ascending order: the merged
    sorted copy, built before the copy's order test is read, is dropped and
    the compacted base is used as its own sorted order by later batches."""
    rng = np.random.default_rng(5)
    ni, nl = H.random_graph(rng, 200, 0.05)
    # new vertices 200..219 hang off nothing; edges among them inserted in
    # ascending order of both endpoints keep every row ascending
    n = 220
    ni = np.concatenate([ni, np.full(n - 200, ni[-1], np.int64)])
    mis = oracle.Graph.from_csr(ni, nl).greedy_mis()
    pairs = [(a, b) for a in range(200, n) for b in range(a + 1, n) if (a * 7 + b) % 3 == 0]
    u = np.array([p[0] for p in pairs], np.int32)
    v = np.array([p[1] for p in pairs], np.int32)
    batches = [(u, v, np.ones(len(u), np.uint8))]
    want, _, og = H.oracle_replay(oracle, ni, nl, mis, batches)
    got, _, _ = H.engine_replay(engine_mod, eng, ni, nl, mis, batches)
    assert np.array_equal(want[-1], got[-1])
    oni, onl = og.compact()
    assert all(np.all(np.diff(onl[oni[x]:oni[x + 1]]) > 0) for x in range(n))  # the case under test
    assert eng.compact() == len(onl)
    eni, enl = eng.export_csr()
    assert np.array_equal(eni, oni) and np.array_equal(enl, onl)
    # later batches search the compacted rows (deletions and insertions of existing pairs)
    more = H.random_batches(rng, oracle, oni, onl, 3, 400, general=True)
    w2, _, _ = H.oracle_replay(oracle, oni, onl, want[-1], more)
    for (bu, bv, bins), w in zip(more, w2):
        eng.process_batch(engine_mod.make_updates(bu, bv, bins))
        assert np.array_equal(eng.get_mis(), w)
    assert eng.check() == (0, 0, 0)


def test_repeated_compactions_reuse_buffers(oracle, engine_mod, eng):
    """One context compacted four times while its base grows a little and
    shrinks more: each compaction's outputs come from the buffers the one
    before released where they fit (the old base, offsets, sorted copy and
    deletion marks), and every compacted base equals the oracle's."""
    rng = np.random.default_rng(23)
    ni, nl = H.random_graph(rng, 3000, 0.004)
    mis = oracle.Graph.from_csr(ni, nl).greedy_mis()
    eng.load_csr(ni, nl)
    eng.set_mis(mis)
    eng.set_rule(True)
    for k in range(4):
        if k % 2 == 0:  # insertions of new pairs: rows grow past their ascending order
            batches = H.random_batches(rng, oracle, ni, nl, 2, 200)
        else:  # deletions of existing entries
            src = np.repeat(np.arange(len(ni) - 1, dtype=np.int32), np.diff(ni))
            pick = [rng.integers(0, len(nl), 1500) for _ in range(2)]
            batches = [(src[p], nl[p], np.zeros(len(p), np.uint8)) for p in pick]
        want, _, og = H.oracle_replay(oracle, ni, nl, mis, batches)
        for (u, v, ins), w in zip(batches, want):
            eng.process_batch(engine_mod.make_updates(u, v, ins))
            assert np.array_equal(eng.get_mis(), w), "phase %d" % k
        oni, onl = og.compact()
        assert eng.compact() == len(onl)
        eni, enl = eng.export_csr()
        assert np.array_equal(eni, oni) and np.array_equal(enl, onl), "compaction %d" % (k + 1)
        ni, nl, mis = oni, onl, want[-1]
    assert eng.check() == (0, 0, 0)


@pytest.mark.parametrize("mode", MODES)
def test_er_config_parity(oracle, engine_mod, eng, mode):
    """Erdos-Renyi n=100,000, 800,000 pairs, seed 42; greedy MIS; 10 batches
    of 2,000 uniform pairs flagged against the loaded graph."""
    ni, nl = engine_mod.gen_er(100000, 800000, 42)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.UNIFORM_PAIRS, ni, nl, 20000, 42)
    g = oracle.Graph.from_csr(ni, nl)
    ins = g.flags(u, v)
    batches = [(u[i:i + 2000], v[i:i + 2000], ins[i:i + 2000]) for i in range(0, 20000, 2000)]
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    got, gst, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, completion=True, mode=mode)
    for b in range(len(batches)):
        assert np.array_equal(want[b], got[b]), "batch %d" % b
        assert checks[b] == (0, 0, 0)
    assert H.counter_mismatch(wst, gst) == ""


def test_determinism(oracle, engine_mod, eng):
    ni, nl = engine_mod.gen_rmat(14, 16, 7)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 8000, 7)
    g = oracle.Graph.from_csr(ni, nl)
    ins = g.flags(u, v)
    batches = [(u[i:i + 2000], v[i:i + 2000], ins[i:i + 2000]) for i in range(0, 8000, 2000)]
    runs = []
    for mode in (1, 0, 1):
        got, _, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, mode=mode)
        runs.append(got)
        assert all(c == (0, 0, 0) for c in checks)
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a, b)
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    for a, b in zip(want, runs[0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", MODES)
def test_nonmaximal_start_parity(oracle, engine_mod, eng, mode):
    """Start from an independent but not maximal set (a third of the greedy
    MIS dropped), run batches without and then with the completion: vertices
    left uncovered before a batch must stay in the rule's reach until the
    state is maximal again."""
    rng = np.random.default_rng(21)
    ni, nl = H.random_graph(rng, 400, 0.02)
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    mis0[np.flatnonzero(mis0)[rng.random(int(mis0.sum())) < 0.33]] = 0
    batches = H.random_batches(rng, oracle, ni, nl, 6, 60, general=True)
    for first_without in (True, False):
        want, got = [], []
        og = oracle.Graph.from_csr(ni, nl)
        m = mis0.copy()
        eng.load_csr(ni, nl)
        eng.set_mis(mis0)
        eng.set_mode(mode)
        for b, (u, v, ins) in enumerate(batches):
            comp = not (first_without and b < 2)
            og.process_batch(m, u, v, ins, completion=comp, literal=True)
            eng.set_rule(comp)
            eng.process_batch(engine_mod.make_updates(u, v, ins))
            assert np.array_equal(eng.get_mis(), m), "batch %d (completion %s)" % (b, comp)
            # the completion only reaches the batch's neighbourhood: vertices
            # uncovered from the start elsewhere stay so, as in the oracle
            conf, unc, bad = eng.check()
            assert (conf, bad) == (0, 0) and unc == og.check(m)[1]


def test_capacity_growth_with_batches_in_flight(oracle, engine_mod, eng):
    """A batch larger than any before regrows the staging buffers while
    earlier batches are still queued: their tickets stay valid and every
    result is the oracle's."""
    rng = np.random.default_rng(17)
    ni, nl = H.random_graph(rng, 600, 0.02)
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    sizes = [50, 40000, 30, 70000]
    batches = []
    for k in sizes:
        u = rng.integers(0, 600, k).astype(np.int32)
        v = rng.integers(0, 600, k).astype(np.int32)
        keep = u != v
        u, v = u[keep], v[keep]
        batches.append((u, v, g.flags(u, v)))
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=True)
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_rule(True)
    eng.set_mode(engine_mod.MODE_QUEUED)
    t0 = eng.submit(engine_mod.make_updates(*batches[0]))
    t1 = eng.submit(engine_mod.make_updates(*batches[1]))   # regrows the staging buffers
    s0 = eng.wait(t0)
    s1 = eng.wait(t1)
    assert s0["updates"] == len(batches[0][0]) and s1["updates"] == len(batches[1][0])
    assert np.array_equal(eng.get_mis(), want[1])
    t2 = eng.submit(engine_mod.make_updates(*batches[2]))
    t3 = eng.submit(engine_mod.make_updates(*batches[3]))
    eng.wait(t2)
    eng.wait(t3)
    assert np.array_equal(eng.get_mis(), want[3])
    assert eng.check() == (0, 0, 0)


def _dense_conflict_stream(oracle, engine_mod, nb=4, bs=8000):
    """A small R-MAT graph (8,192 vertices) with batches about as large as
    its vertex count, endpoints in its lowest-degree quarter: far more
    speculative rounds per update than a first estimate allows."""
    ni, nl = engine_mod.gen_rmat(13, 16, 21)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.LOW_DEGREE, ni, nl, nb * bs, 21)
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    batches = [(u[i:i + bs], v[i:i + bs], ins[i:i + bs]) for i in range(0, nb * bs, bs)]
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    return ni, nl, mis0, batches, want


@pytest.mark.parametrize("driver", ["wait", "poll", "submit"])
def test_rounds_beyond_first_estimate(oracle, engine_mod, driver):
    """Queued batches on a fresh context (no round history) that need many
    more rounds than were queued: the tail halts and the rest of the rounds
    are queued from the host, by wait, by a polling loop, or by the next
    submit of the same graph; every MIS is the oracle's."""
    ni, nl, mis0, batches, want = _dense_conflict_stream(oracle, engine_mod)
    with engine_mod.Engine(0) as e:
        e.load_csr(ni, nl)
        e.set_mis(mis0)
        e.set_mode(engine_mod.MODE_QUEUED)
        ups = [engine_mod.make_updates(*b) for b in batches]
        if driver == "submit":
            t = [e.submit(ups[0]), e.submit(ups[1])]   # the second submit settles the first batch
            st = [e.wait(t[0]), e.wait(t[1])]
            assert np.array_equal(e.get_mis(), want[1])
            rest = range(2, len(batches))
        else:
            t0 = e.submit(ups[0])
            if driver == "poll":
                polls = 0
                while not e.poll(t0):
                    polls += 1
                assert polls >= 0
            st = [e.wait(t0)]
            assert np.array_equal(e.get_mis(), want[0])
            rest = range(1, len(batches))
        # the first batch was queued with ceil(8000 / 400) + 3 grid rounds
        assert st[0]["rounds"] > 23
        for b in rest:
            e.wait(e.submit(ups[b]))
            assert np.array_equal(e.get_mis(), want[b]), "batch %d" % b
        assert e.check() == (0, 0, 0)


@pytest.mark.parametrize("mode", MODES)
def test_unsorted_rows_parity(oracle, engine_mod, eng, mode):
    """Rows in arbitrary order (a graph converted from a file in file order):
    the order is kept (it decides which vertex an eviction admits) and
    adjacency tests go through the sorted row copy."""
    rng = np.random.default_rng(23)
    ni, nl = H.random_graph(rng, 500, 0.03)
    nl = nl.copy()
    for x in range(500):
        rng.shuffle(nl[ni[x]:ni[x + 1]])
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 5, 100, general=True)
    want, _, og = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=True)
    got, _, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, mode=mode)
    for b in range(len(batches)):
        assert np.array_equal(want[b], got[b]), "batch %d" % b
        assert checks[b] == (0, 0, 0)
    for v in range(0, 500, 11):
        assert np.array_equal(eng.neighbors(v), og.neighbors(v))
    dup = nl.copy()
    dup[ni[3] + 1] = dup[ni[3]]  # a row listing a neighbour twice is refused
    with pytest.raises(engine_mod.MisfitError):
        eng.load_csr(ni, dup)


def _star_and_friends():
    """(name, nindex, nlist, batches) of small corner-case graphs."""
    cases = []
    # no edges at all: a batch inserts some, the next deletes one again
    ni = np.zeros(9, np.int64)
    nl = np.zeros(0, np.int32)
    cases.append(("empty graph", ni, nl, [
        (np.array([0, 1, 2, 5], np.int32), np.array([1, 2, 3, 6], np.int32), np.ones(4, np.uint8)),
        (np.array([1], np.int32), np.array([2], np.int32), np.zeros(1, np.uint8))]))
    # one vertex: only an empty batch is possible
    cases.append(("one vertex", np.zeros(2, np.int64), np.zeros(0, np.int32),
                  [(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8))]))
    # a star (hub 0 with 1,200 leaves) plus isolated vertices 1201..1210:
    # delete spokes, connect leaves to each other and to isolated vertices
    n = 1211
    rows = [list(range(1, 1201))] + [[0] for _ in range(1200)] + [[] for _ in range(10)]
    ni = np.zeros(n + 1, np.int64)
    for x in range(n):
        ni[x + 1] = ni[x] + len(rows[x])
    nl = np.array([y for r in rows for y in r], np.int32)
    rng = np.random.default_rng(19)
    dels = rng.choice(np.arange(1, 1201), 300, replace=False).astype(np.int32)
    a = rng.integers(1, 1211, 400).astype(np.int32)
    b = rng.integers(1, 1211, 400).astype(np.int32)
    keep = a != b
    a, b = a[keep], b[keep]
    cases.append(("star", ni, nl, [
        (np.zeros(300, np.int32), dels, np.zeros(300, np.uint8)),       # hub loses 300 spokes
        (a, b, np.ones(len(a), np.uint8)),                               # leaf-leaf and isolated links
        (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8)),  # an empty batch
        (np.zeros(50, np.int32), dels[:50], np.ones(50, np.uint8))]))    # 50 spokes come back
    return cases


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", range(3))
def test_corner_case_graphs(oracle, engine_mod, eng, mode, case):
    """A graph without edges, a one-vertex graph, and a star with isolated
    vertices, with empty batches among the others: the MIS after every batch
    is the oracle's (flags as given: applied as insert-if-absent /
    delete-if-present), and the full check passes."""
    name, ni, nl, batches = _star_and_friends()[case]
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches)
    got, stats, checks = H.engine_replay(engine_mod, eng, ni, nl, mis0, batches, mode=mode)
    for k, (w, g) in enumerate(zip(want, got)):
        assert np.array_equal(w, g), "%s: batch %d" % (name, k)
    assert all(c == (0, 0, 0) for c in checks), name
    for (u, _, _), st in zip(batches, stats):
        assert st["updates"] == len(u)


@pytest.mark.parametrize("mode", MODES)
def test_tagged_scan_equals_full_scan(oracle, engine_mod, mode, monkeypatch):
    """With a maximal MIS before every batch the may-uncover scan tests only
    the vertices the k-count tagged; forcing the scan over every vertex
    (MISFIT_PU_SCAN=full, read when a context is created) must give the same
    MIS after every batch, and both the oracle's."""
    ni, nl = engine_mod.gen_rmat(15, 16, 11)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 12000, 11)
    g = oracle.Graph.from_csr(ni, nl)
    ins = g.flags(u, v)
    batches = [(u[i:i + 3000], v[i:i + 3000], ins[i:i + 3000]) for i in range(0, 12000, 3000)]
    runs = []
    for scan in ("", "full"):
        if scan:
            monkeypatch.setenv("MISFIT_PU_SCAN", scan)
        else:
            monkeypatch.delenv("MISFIT_PU_SCAN", raising=False)
        with engine_mod.Engine(0) as e:
            got, stats, checks = H.engine_replay(engine_mod, e, ni, nl, mis0, batches, mode=mode)
        assert all(c == (0, 0, 0) for c in checks)
        runs.append((got, [s["maybe_uncovered"] for s in stats]))
    for a, b in zip(runs[0][0], runs[1][0]):
        assert np.array_equal(a, b)
    assert runs[0][1] == runs[1][1]  # the same PU set size both ways
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    for a, b in zip(want, runs[0][0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("knobs", [{"MISFIT_ROUND_THREADS": "128"}, {"MISFIT_ROUND_THREADS": "256"},
                                   {"MISFIT_PU_DENSE_FRACTION": "0"},
                                   {"MISFIT_ROUND_THREADS": "256", "MISFIT_ROUND_WG_PER_CU": "1"}])
def test_tuning_knobs_keep_results(oracle, engine_mod, mode, knobs, monkeypatch):
    """The tuner's knobs change how the work is laid out, never the result:
    the round kernels in workgroups of 128 or 256 threads (the default is 512; windows
    follow the grid's waves), and the may-uncover scan switched to every
    vertex for every batch, each give the oracle's MIS after every batch,
    with the device check passing."""
    for k in ("MISFIT_ROUND_THREADS", "MISFIT_PU_DENSE_FRACTION", "MISFIT_ROUND_WG_PER_CU", "MISFIT_PU_SCAN"):
        monkeypatch.delenv(k, raising=False)
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    ni, nl = engine_mod.gen_rmat(15, 16, 13)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 12000, 13)
    g = oracle.Graph.from_csr(ni, nl)
    ins = g.flags(u, v)
    sizes = [500, 3000, 6000, 2500]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    batches = [(u[a:b], v[a:b], ins[a:b]) for a, b in zip(offs[:-1], offs[1:])]
    with engine_mod.Engine(0) as e:
        got, stats, checks = H.engine_replay(engine_mod, e, ni, nl, mis0, batches, mode=mode)
    assert all(c == (0, 0, 0) for c in checks)
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    for a, b in zip(want, got):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("knobs", [{"MISFIT_GRID_ROUNDS": "200"}, {}, {"MISFIT_GRID_ROUNDS": "1"}])
def test_rule_counters_match_oracle(oracle, engine_mod, knobs, monkeypatch):
    """Queued batches report the oracle's rule counters (active insertions,
    evictions, admissions on both branches, completion additions) whether
    their edges are committed by grid-wide rounds (200 fixed rounds: every
    batch ends in the grid, the rest run empty), by the planned rounds, or
    mostly by the one-workgroup tail (one grid round), with the MIS equal
    after every batch."""
    monkeypatch.delenv("MISFIT_GRID_ROUNDS", raising=False)
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    ni, nl = engine_mod.gen_rmat(15, 16, 17)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 9000, 17)
    g = oracle.Graph.from_csr(ni, nl)
    ins = g.flags(u, v)
    sizes = [400, 2600, 5000, 1000]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    batches = [(u[a:b], v[a:b], ins[a:b]) for a, b in zip(offs[:-1], offs[1:])]
    with engine_mod.Engine(0) as e:
        got, gst, checks = H.engine_replay(engine_mod, e, ni, nl, mis0, batches, mode=1)
    assert all(c == (0, 0, 0) for c in checks)
    want, wst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    for a, b in zip(want, got):
        assert np.array_equal(a, b)
    assert sum(s["evictions"] for s in wst) > 0 and sum(s["admit_del"] for s in wst) > 0
    assert H.counter_mismatch(wst, gst) == ""


@pytest.mark.parametrize("mode", MODES)
def test_device_times_and_timer(engine_mod, eng, mode):
    """Per-batch device times come from the device clock (the prologue is part
    of the batch), and the event timer brackets queued work."""
    ni, nl = engine_mod.gen_rmat(14, 16, 3)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 6000, 3)
    eng.load_csr(ni, nl)
    ups = engine_mod.make_updates(u, v, eng.flags(u, v))
    eng.set_mis(mis0)
    eng.set_mode(mode)
    eng.sync()
    eng.timer_begin()
    tickets = [eng.submit(ups[i:i + 2000]) for i in range(0, 6000, 2000)]
    stats = [eng.wait(t) for t in tickets]
    region = eng.timer_end()
    for s in stats:
        assert 0.0 < s["prologue_ms"] <= s["device_ms"] < 1000.0
    # the region holds every batch's device time
    assert region >= 0.9 * sum(s["device_ms"] for s in stats)
    assert eng.check() == (0, 0, 0)


@pytest.mark.parametrize("mode", MODES)
def test_staged_batches_parity(oracle, engine_mod, eng, mode):
    """Batches uploaded ahead of submission give the oracle's MIS after every
    batch, two in flight; an empty batch stages; a staged batch is refused by
    another context; release frees it after the work that reads it."""
    rng = np.random.default_rng(31)
    n = 500
    ni, nl = H.random_graph(rng, n, 0.02)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 6, 120, general=True)
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=True)
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_rule(True)
    eng.set_mode(mode)
    hs = [eng.stage(engine_mod.make_updates(*b)) for b in batches]
    empty = eng.stage(engine_mod.make_updates(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint8)))
    tickets = [eng.submit_staged(hs[0])]
    for b in range(len(batches)):
        if b + 1 < len(batches):
            tickets.append(eng.submit_staged(hs[b + 1]))
        st = eng.wait(tickets[b])
        assert st["updates"] == len(batches[b][0])
        if b + 1 == len(batches):
            assert np.array_equal(eng.get_mis(), want[b])
    assert eng.wait(eng.submit_staged(empty))["updates"] == 0
    assert np.array_equal(eng.get_mis(), want[-1])
    assert eng.check() == (0, 0, 0)
    with engine_mod.Engine(0) as other:
        other.load_csr(ni, nl)
        with pytest.raises(engine_mod.MisfitError):
            other.submit_staged(hs[0])
    # one at a time: the MIS after every batch
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    for b in range(len(batches)):
        eng.wait(eng.submit_staged(hs[b]))
        assert np.array_equal(eng.get_mis(), want[b]), "batch %d differs" % b
    for h in hs + [empty]:
        eng.release_staged(h)
    with pytest.raises(engine_mod.MisfitError):
        eng.submit_staged(hs[0])


@pytest.mark.parametrize("mode", MODES)
def test_invalid_updates_refused_without_effect(oracle, engine_mod, eng, mode):
    """Endpoints out of range, negative or equal are refused by submit, stage
    and process_batch with MISFIT_E_ARG naming the update, before anything
    reaches the device: the MIS and the graph stay as they were, and the
    next valid batch gives the oracle's result."""
    rng = np.random.default_rng(41)
    n = 200
    ni, nl = H.random_graph(rng, n, 0.04)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 2, 60, general=True)
    want, _, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=True)
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_rule(True)
    eng.set_mode(mode)
    eng.wait(eng.submit(engine_mod.make_updates(*batches[0])))
    assert np.array_equal(eng.get_mis(), want[0])
    size0 = eng.graph_size()
    u, v, ins = (np.array(a) for a in batches[1])
    for bad_u, bad_v in ((n, 0), (-1, 3), (7, 7), (0, n + 5)):
        bu, bv = u.copy(), v.copy()
        bu[17], bv[17] = bad_u, bad_v
        ups = engine_mod.make_updates(bu, bv, ins)
        for call in (eng.submit, eng.stage, eng.process_batch):
            with pytest.raises(engine_mod.MisfitError) as ei:
                call(ups)
            assert ei.value.code == -1 and "update 17" in str(ei.value)
    assert np.array_equal(eng.get_mis(), want[0])
    assert eng.graph_size() == size0
    eng.wait(eng.submit(engine_mod.make_updates(u, v, ins)))
    assert np.array_equal(eng.get_mis(), want[1])
    assert eng.check() == (0, 0, 0)


def test_failed_capacity_growth_leaves_the_context_usable(oracle, engine_mod, monkeypatch):
    """Fault injection (MISFIT_TEST_FAIL_GROWTH=1: every capacity growth of
    the context fails): a batch larger than the batch buffers fails with
    MISFIT_E_NOMEM and changes nothing; the next, smaller batches, the flags
    and a compaction work on the old buffers, and the MIS equals the oracle's
    over the batches that ran."""
    monkeypatch.setenv("MISFIT_TEST_FAIL_GROWTH", "1")
    ni, nl = engine_mod.gen_er(1 << 13, 8 << 13, 71)
    mis0 = engine_mod.greedy_mis(ni, nl)
    u, v = engine_mod.gen_updates(engine_mod.MIXED_EXISTING, ni, nl, 24000, 71)
    ins = oracle.Graph.from_csr(ni, nl).flags(u, v)
    big = (u[:20000], v[:20000], ins[:20000])          # past the first 16,384-update buffers
    small = [(u[a:a + 2000], v[a:a + 2000], ins[a:a + 2000]) for a in (20000, 22000)]
    with engine_mod.Engine(0) as eng:
        eng.load_csr(ni, nl)
        eng.set_mis(mis0)
        with pytest.raises(engine_mod.MisfitError) as ei:
            eng.process_batch(engine_mod.make_updates(*big))
        assert ei.value.code == -4
        assert np.array_equal(eng.get_mis(), mis0)
        assert np.array_equal(eng.flags(u[20000:], v[20000:]), ins[20000:])
        for b in small:
            eng.process_batch(engine_mod.make_updates(*b))
        assert eng.check() == (0, 0, 0)
        want, _, og = H.oracle_replay(oracle, ni, nl, mis0, small, completion=True, literal=False)
        assert np.array_equal(eng.get_mis(), want[-1])
        eng.compact()
        cni, cnl = og.compact()
        gni, gnl = eng.export_csr()
        assert np.array_equal(gni, cni) and np.array_equal(gnl, cnl)
