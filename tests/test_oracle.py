"""Pins the CPU oracle (oracle/misfit_oracle.c) before anything is compared
against it.

  * reference fixtures: datasets/sample of the reference (copied into
    tests/golden/sample) -- header, greedy MIS, flags, and the known answer of
    the literal rule re-derived here and committed as tests/golden/sample_known_answer.json;
  * the reference's own host classes (core/FlashGraph.h, core/flash_compactor.h)
    compiled into oracle/_ref: overlay order, adjacency, degrees and the
    compacted file must match byte for byte;
  * hand-computed known answers for every branch of the rule
    (src/misfit/misfit.cu:283-347) and for the completion pass;
  * an independent brute-force restatement on tiny random inputs, and the two
    oracle engines (literal rows vs lazy) against each other;
  * invariants: the literal rule keeps independence; completion keeps maximality.
"""
import json
import os
import struct

import numpy as np
import pytest

from tests import helpers as H
from tests.conftest import GOLDEN, SAMPLE


# ------------------------------------------------------------ sample fixtures
def test_sample_header_and_shape(oracle):
    raw = open(os.path.join(SAMPLE, "sample.egr"), "rb").read()
    n, sent, m = struct.unpack("<iiq", raw[:16])
    assert (n, sent, m) == (2000, -1, 16000)
    assert len(raw) == 16 + 2001 * 8 + 16000 * 4 == 80024
    g = oracle.Graph.load_egr(os.path.join(SAMPLE, "sample.egr"))
    ni, nl = g.base_csr()
    assert g.wide and g.n == 2000 and g.m == 16000
    deg = np.diff(ni)
    assert deg.min() == 0 and deg.max() == 19 and deg.mean() == 8
    src = np.repeat(np.arange(2000), deg)
    assert not np.any(src == nl)
    for x in range(2000):
        r = nl[ni[x]:ni[x + 1]]
        assert np.all(np.diff(r) > 0)
    pairs = set(zip(src.tolist(), nl.tolist()))
    assert all((b, a) in pairs for (a, b) in pairs)


def test_sample_mis_is_greedy(oracle):
    g = oracle.Graph.load_egr(os.path.join(SAMPLE, "sample.egr"))
    mis, card = oracle.read_mis(os.path.join(SAMPLE, "sample_mis.txt"), g.n)
    assert card == 533 and int(mis.sum()) == 533
    assert np.array_equal(g.greedy_mis(), mis)
    assert g.check(mis) == (0, 0)


def test_sample_batches_are_distinct_deletions(oracle):
    g, mis, batches = H.sample_inputs(oracle, SAMPLE)
    assert [len(b[0]) for b in batches] == [50] * 5
    allp = set()
    for (u, v, ins) in batches:
        assert not ins.any()  # every pair is an existing edge: all deletions
        for a, b in zip(u.tolist(), v.tolist()):
            allp.add((min(a, b), max(a, b)))
    assert len(allp) == 250


def test_sample_known_answer(oracle):
    """Literal rule over the 5 sample batches (re-derived; see the JSON)."""
    golden = json.load(open(os.path.join(GOLDEN, "sample_known_answer.json")))
    g, mis0, batches = H.sample_inputs(oracle, SAMPLE)
    ni, nl = g.base_csr()
    for literal in (True, False):
        for completion in (False, True):
            key = "completion" if completion else "literal"
            gg = oracle.Graph.from_csr(ni, nl)
            m = np.array(mis0, np.uint8)
            adm, adds, unc = [], [], []
            for (u, v, ins) in batches:
                st = gg.process_batch(m, u, v, ins, completion=completion, literal=literal)
                adm.append(st["admit_del"] + st["admit_ins"])
                adds.append(st["completion_adds"])
                conf, un = gg.check(m)
                assert conf == 0
                unc.append(un)
            want = golden[key]
            assert adm == want["admissions"], key
            assert adds == want["completion_adds"], key
            assert unc == want["uncovered"], key
            assert int(m.sum()) == want["final_card"], key
            assert [int(i) + 1 for i in np.nonzero(m)[0]] == want["final_ids"], key
            assert len(gg.compact()[1]) == want["compacted_entries"], key
    # the survey's independently derived numbers
    assert golden["literal"]["admissions"] == [0, 0, 0, 0, 0]
    assert golden["literal"]["uncovered"] == [4, 9, 12, 15, 19]
    assert golden["literal"]["final_card"] == 533
    assert golden["literal"]["compacted_entries"] == 15500
    assert golden["completion"]["uncovered"] == [0, 0, 0, 0, 0]


# ------------------------------------------------------------ reference classes
needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libmisfit_ref.so")),
    reason="oracle/_ref not built (reference checkout absent)")


@needs_ref
def test_overlay_matches_reference_classes(oracle, tmp_path):
    """Random insert/remove sequences: neighbour order, adjacency, degree and
    operation count equal the reference's HybridGraph."""
    rng = np.random.default_rng(3)
    for trial in range(3):
        ni, nl = H.random_graph(rng, 60, 0.1)
        path = str(tmp_path / ("g%d.egr" % trial))
        _write_wide(path, ni, nl)
        ref = oracle.RefGraph(path)
        mine = oracle.Graph.load_egr(path)
        for step in range(600):
            u, v = (int(x) for x in rng.integers(0, 60, 2))
            if u == v:
                continue
            a1, a2 = ref.is_adjacent(u, v), mine.is_adjacent(u, v)
            assert a1 == a2
            if rng.random() < 0.5:
                if not a1:
                    ref.insert_edge(u, v)
                    mine.insert_edge(u, v)
            elif a1:
                ref.remove_edge(u, v)
                mine.remove_edge(u, v)
        assert ref.delta_ops == mine.delta_ops
        for x in range(60):
            assert np.array_equal(ref.neighbors(x), mine.neighbors(x)), x
            assert ref.degree(x) == mine.degree(x)
        ref.close()


@needs_ref
def test_compaction_matches_reference_file(oracle, tmp_path):
    """The oracle's compacted CSR, written with the wide layout, is byte-identical
    to the file the reference's compact_flash_graph writes."""
    rng = np.random.default_rng(4)
    g0, mis0, batches = H.sample_inputs(oracle, SAMPLE)
    ni, nl = g0.base_csr()
    p_ref = str(tmp_path / "ref.egr")
    _write_wide(p_ref, ni, nl)
    ref = oracle.RefGraph(p_ref)
    mine = oracle.Graph.from_csr(ni, nl)
    for (u, v, ins) in batches:
        mine.apply_batch(u, v, ins)
        for a, b, f in zip(u.tolist(), v.tolist(), ins.tolist()):
            adj = ref.is_adjacent(a, b)
            if f and not adj:
                ref.insert_edge(a, b)
            elif not f and adj:
                ref.remove_edge(a, b)
    for _ in range(200):  # some insertions too
        a, b = (int(x) for x in rng.integers(0, 2000, 2))
        if a != b and not ref.is_adjacent(a, b):
            ref.insert_edge(a, b)
            mine.insert_edge(a, b)
    ref.compact_to(p_ref)
    ref.close()
    cni, cnl = mine.compact()
    p_mine = str(tmp_path / "mine.egr")
    _write_wide(p_mine, cni, cnl)
    assert open(p_ref, "rb").read() == open(p_mine, "rb").read()


@needs_ref
def test_reference_batch_apply(oracle, tmp_path):
    """ref_apply (the reference's overlay update of a batch, misfit.cu:708-714)
    gives the same rows as applying the batch operation by operation."""
    rng = np.random.default_rng(8)
    ni, nl = H.random_graph(rng, 80, 0.08)
    p1, p2 = str(tmp_path / "a.egr"), str(tmp_path / "b.egr")
    _write_wide(p1, ni, nl)
    _write_wide(p2, ni, nl)
    a, b = oracle.RefGraph(p1), oracle.RefGraph(p2)
    u = rng.integers(0, 80, 3000).astype(np.int32)
    v = rng.integers(0, 80, 3000).astype(np.int32)
    keep = u != v
    u, v = u[keep], v[keep]
    ins = (rng.random(len(u)) < 0.5).astype(np.uint8)
    a.apply(u, v, ins)
    # without the adjacency tests on a stream whose updates all change the graph
    ni2, nl2 = H.random_graph(rng, 80, 0.08)
    p3, p4 = str(tmp_path / "c.egr"), str(tmp_path / "d.egr")
    _write_wide(p3, ni2, nl2)
    _write_wide(p4, ni2, nl2)
    c_, d_ = oracle.RefGraph(p3), oracle.RefGraph(p4)
    pairs = [(x, y) for x in range(80) for y in range(x + 1, 80)]
    pick = rng.permutation(len(pairs))[:600]
    eu = np.array([pairs[k][0] for k in pick], np.int32)
    ev = np.array([pairs[k][1] for k in pick], np.int32)
    eins = np.array([0 if c_.is_adjacent(int(x), int(y)) else 1 for x, y in zip(eu, ev)], np.uint8)
    c_.apply(eu, ev, eins, check=False)
    d_.apply(eu, ev, eins, check=True)
    assert c_.delta_ops == d_.delta_ops
    for x in range(80):
        assert np.array_equal(c_.neighbors(x), d_.neighbors(x))
    c_.close()
    d_.close()
    for x, y, f in zip(u.tolist(), v.tolist(), ins.tolist()):
        adj = b.is_adjacent(x, y)
        if f and not adj:
            b.insert_edge(x, y)
        elif not f and adj:
            b.remove_edge(x, y)
    assert a.delta_ops == b.delta_ops
    for x in range(80):
        assert np.array_equal(a.neighbors(x), b.neighbors(x))
    a.close()
    b.close()


def _write_wide(path, ni, nl):
    with open(path, "wb") as f:
        f.write(struct.pack("<iiq", len(ni) - 1, -1, len(nl)))
        f.write(np.asarray(ni, np.int64).tobytes())
        f.write(np.asarray(nl, np.int32).tobytes())


# ------------------------------------------------------------ hand-computed KATs
def _csr(n, edges, order=None):
    """CSR from an undirected edge list; `order` optionally fixes row orders."""
    rows = [[] for _ in range(n)]
    for a, b in edges:
        rows[a].append(b)
        rows[b].append(a)
    for x in range(n):
        rows[x] = sorted(rows[x])
    if order:
        for x, r in order.items():
            assert sorted(r) == rows[x]
            rows[x] = list(r)
    ni = np.zeros(n + 1, np.int64)
    for x in range(n):
        ni[x + 1] = ni[x] + len(rows[x])
    nl = np.array([y for r in rows for y in r], np.int32)
    return ni, nl


def _run(oracle, ni, nl, mis_ids, batch, completion, literal=True):
    g = oracle.Graph.from_csr(ni, nl)
    mis = np.zeros(len(ni) - 1, np.uint8)
    mis[list(mis_ids)] = 1
    u = np.array([p[0] for p in batch], np.int32)
    v = np.array([p[1] for p in batch], np.int32)
    ins = g.flags(u, v)
    g.process_batch(mis, u, v, ins, completion=completion, literal=literal)
    return sorted(np.nonzero(mis)[0].tolist())


# Each case was worked out by hand from the rule; the comment names the
# mistake it catches.
KATS = [
    # name, n, edges, row order, initial MIS, batch, literal result, completion result
    # (catches: evicting min(u,v); scanning past the first admissible vertex)
    # eviction of max(u,v)=1, then the first admissible in BFS order: 2
    ("evict_max_admit_first", 6, [(1, 2), (1, 3), (3, 5)], None, {0, 1, 4, 5}, [(0, 1)],
     [0, 2, 4, 5], [0, 2, 4, 5]),
    # row(1) = [1, 3, 2, 0]: 3 precedes 2 in base order, so 3 is admitted
    # (catches: sorted or hashed rows instead of BFS discovery order)
    ("bfs_order_not_sorted", 5, [(1, 2), (1, 3)], {1: [3, 2]}, {0, 1, 4}, [(0, 1)],
     [0, 3, 4], [0, 2, 3, 4]),
    # deletion: row(1) = [1] only, no member other than 0 -> 1 joins
    ("delete_admits", 3, [(0, 1)], None, {0, 2}, [(0, 1)], [0, 1, 2], [0, 1, 2]),
    # deletion: 3 sits two hops from 1 -> blocked (catches: 1-hop rows)
    ("delete_blocked_two_hops", 4, [(0, 1), (1, 2), (2, 3)], None, {0, 3}, [(0, 1)],
     [0, 3], [0, 1, 3]),
    # deletion in a triangle: 0 is two hops away via 2 but is the excluded end
    # (catches: not excluding the other endpoint)
    ("delete_excludes_other_end", 3, [(0, 1), (1, 2), (0, 2)], None, {0}, [(0, 1)],
     [0, 1], [0, 1]),
    # 2 is an endpoint (seed) so its row is 2-hop and holds 0; 3 is admitted
    # (catches: 1-hop rows for seed candidates)
    ("seed_candidate_two_hop", 8, [(1, 2), (1, 3), (2, 6), (6, 7)], None, {0, 1, 4, 7},
     [(2, 6), (0, 1)], [0, 3, 4, 7], [0, 2, 3, 4, 7]),
    # 5 is free but has no row (not next to an endpoint): skipped
    # (catches: admitting vertices without a row)
    ("no_row_skipped", 7, [(1, 2), (2, 5), (2, 6)], None, {0, 1, 3, 4, 6}, [(0, 1)],
     [0, 3, 4, 6], [0, 3, 4, 6]),
]


@pytest.mark.parametrize("kat", KATS, ids=[k[0] for k in KATS])
def test_rule_known_answers(oracle, kat):
    name, n, edges, order, mis, batch, want_lit, want_comp = kat
    ni, nl = _csr(n, edges, order)
    for literal_engine in (True, False):
        assert _run(oracle, ni, nl, mis, batch, False, literal_engine) == want_lit, name
        assert _run(oracle, ni, nl, mis, batch, True, literal_engine) == want_comp, name


# ------------------------------------------------------------ brute force
@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(8, 40))
    ni, nl = H.random_graph(rng, n, float(rng.uniform(0.05, 0.3)))
    adj0 = [nl[ni[x]:ni[x + 1]].tolist() for x in range(n)]
    g = oracle.Graph.from_csr(ni, nl)
    mis0 = g.greedy_mis()
    if seed % 3 == 0:  # start from a non-maximal set too
        mis0[np.nonzero(mis0)[0][::3]] = 0
    general = seed % 2 == 1
    batches = H.random_batches(rng, oracle, ni, nl, 3, 12, general=general)
    for completion in (False, True):
        for literal in (True, False):
            gg = oracle.Graph.from_csr(ni, nl)
            m = mis0.copy()
            ost = [gg.process_batch(m, u, v, ins, completion=completion, literal=literal) for (u, v, ins) in batches]
            bm, rows_state, bst = _brute_multi(n, adj0, [int(x) for x in mis0], batches, completion)
            assert [int(x) for x in m] == bm, (seed, completion, literal)
            for x in range(n):
                assert gg.neighbors(x).tolist() == rows_state[x]
            # the rule's counters, batch by batch
            assert H.counter_mismatch(bst, ost) == "", (seed, completion, literal)


def _brute_multi(n, adj0, mis, batches, completion):
    """Independent restatement on Python lists, batch after batch: the overlay
    before a batch is rebuilt from the base plus every earlier update, rows
    are explicit BFS lists, the rule runs in batch order, then completion."""
    hist, counts = [], []
    for (u, v, ins) in batches:
        cur = list(zip(u.tolist(), v.tolist(), ins.tolist()))
        # graph state before this batch = base plus earlier updates; realise it
        # by replaying them as plain graph updates, then the batch with the rule
        state_adj, removed_added = _materialise(n, adj0, hist)
        mis, rows, cnt = _brute_from_state(n, adj0, state_adj, removed_added, mis, cur, completion)
        counts.append(cnt)
        hist.extend(cur)
    return mis, rows, counts


def _materialise(n, adj0, hist):
    removed = [set() for _ in range(n)]
    added = [[] for _ in range(n)]
    base = adj0

    def adjacent(a, b):
        if b in added[a]:
            return True
        if b in removed[a]:
            return False
        return b in base[a]

    for (a, b, ins) in hist:
        adj = adjacent(a, b)
        if ins and not adj:
            if b in removed[a]:
                removed[a].discard(b)
                removed[b].discard(a)
            else:
                added[a].append(b)
                added[b].append(a)
        elif not ins and adj:
            if b in added[a]:
                added[a].remove(b)
                added[b].remove(a)
            else:
                removed[a].add(b)
                removed[b].add(a)
    return None, (removed, added)


def _brute_from_state(n, adj0, _unused, ra, mis, batch, completion):
    removed, added = [set(s) for s in ra[0]], [list(a) for a in ra[1]]
    # run the single-batch brute force with a pre-populated overlay
    base = adj0

    def adjacent(a, b):
        if b in added[a]:
            return True
        if b in removed[a]:
            return False
        return b in base[a]

    def nbrs(x):
        return [y for y in base[x] if y not in removed[x]] + added[x]

    for (a, b, ins) in batch:
        adj = adjacent(a, b)
        if ins and not adj:
            if b in removed[a]:
                removed[a].discard(b)
                removed[b].discard(a)
            else:
                added[a].append(b)
                added[b].append(a)
        elif not ins and adj:
            if b in added[a]:
                added[a].remove(b)
                added[b].remove(a)
            else:
                removed[a].add(b)
                removed[b].add(a)
    seeds = []
    for (a, b, _) in batch:
        for x in (a, b):
            if x not in seeds:
                seeds.append(x)
    rows = {}
    for s in seeds:
        row, seen = [s], {s}
        for y in nbrs(s):
            if y not in seen:
                seen.add(y)
                row.append(y)
        for x in list(row[1:]):
            for y in nbrs(x):
                if y not in seen:
                    seen.add(y)
                    row.append(y)
        rows[s] = row
    for s in seeds:
        for x in nbrs(s):
            if x not in rows:
                rows[x] = nbrs(x)
    mis = list(mis)
    evicted = []
    cnt = dict.fromkeys(H.RULE_COUNTERS, 0)
    for (a, b, ins) in batch:
        if ins:
            if mis[a] and mis[b]:
                cnt["active_ins"] += 1
                w = max(a, b)
                mis[w] = 0
                evicted.append(w)
                cnt["evictions"] += 1
                for x in rows.get(w, []):
                    if mis[x] or x not in rows:
                        continue
                    if not any(mis[y] for y in rows[x]):
                        mis[x] = 1
                        cnt["admit_ins"] += 1
                        break
        else:
            if mis[a] and not mis[b]:
                cnt["active_del"] += 1
                if not any(mis[y] for y in rows[b] if y != a):
                    mis[b] = 1
                    cnt["admit_del"] += 1
            elif not mis[a] and mis[b]:
                cnt["active_del"] += 1
                if not any(mis[y] for y in rows[a] if y != b):
                    mis[a] = 1
                    cnt["admit_del"] += 1
    if completion:
        C = set(seeds) | set(evicted)
        for w in evicted:
            C |= set(nbrs(w))
        for x in sorted(C):
            if not mis[x] and not any(mis[y] for y in nbrs(x)):
                mis[x] = 1
                cnt["completion_adds"] += 1
    return mis, [nbrs(x) for x in range(n)], cnt


# ------------------------------------------------------------ engines + invariants
@pytest.mark.parametrize("seed", range(6))
def test_literal_and_lazy_agree(oracle, seed):
    rng = np.random.default_rng(200 + seed)
    ni, nl = H.random_graph(rng, 250, 0.04)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 5, 60, general=bool(seed % 2))
    for completion in (False, True):
        a, ast, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion, literal=True)
        b, bst, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion, literal=False)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert H.counter_mismatch(ast, bst) == ""


def test_counter_check_catches_a_planted_miscount(oracle):
    """The per-batch counter comparison names the batch and the counter of a
    single miscount planted in a copy of the oracle's stats."""
    rng = np.random.default_rng(77)
    ni, nl = H.random_graph(rng, 250, 0.04)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 4, 60, general=False)
    _, st, _ = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=True, literal=False)
    assert H.counter_mismatch(st, [dict(s) for s in st]) == ""
    assert sum(s["evictions"] for s in st) > 0
    for k in H.RULE_COUNTERS:
        planted = [dict(s) for s in st]
        planted[2][k] += 1
        msg = H.counter_mismatch(st, planted)
        assert msg.startswith("batch 2 %s:" % k), msg
    assert H.counter_mismatch(st, st[:-1]) != ""


@pytest.mark.parametrize("seed", range(4))
def test_invariants(oracle, seed):
    rng = np.random.default_rng(300 + seed)
    ni, nl = H.random_graph(rng, 300, 0.03)
    mis0 = oracle.Graph.from_csr(ni, nl).greedy_mis()
    batches = H.random_batches(rng, oracle, ni, nl, 6, 80, general=False)
    for completion in (False, True):
        g = oracle.Graph.from_csr(ni, nl)
        m = mis0.copy()
        for (u, v, ins) in batches:
            g.process_batch(m, u, v, ins, completion=completion)
            conf, unc = g.check(m)
            assert conf == 0  # the literal rule keeps independence
            if completion:
                assert unc == 0


# ------------------------------------------------------------ file parsing
def test_batch_file_parse_semantics(oracle, tmp_path):
    """core/cli_setup.h:66-82: 1-indexed, two ints per line via istream,
    out-of-range and self-loop pairs dropped, flag = not adjacent."""
    g = oracle.Graph.from_csr(*_csr(5, [(0, 1), (1, 2)]))
    p = tmp_path / "1.mtx"
    p.write_text("%%MatrixMarket header\n5 5 2\n1 2\n2 3 extra\n3 1\n0 4\n6 1\n4 4\n+2  5\n12abc 3\n2 4xyz\n\n1\n")
    u, v, ins = g.parse_batch_file(str(p))
    assert list(zip(u.tolist(), v.tolist(), ins.tolist())) == [
        (0, 1, 0), (1, 2, 0), (2, 0, 1), (1, 4, 1), (1, 3, 1)]


def test_batch_files_sorted_numerically(oracle, tmp_path):
    for k in (10, 2, 1):
        (tmp_path / ("%d.mtx" % k)).write_text("1 2\n")
    names = [os.path.basename(p) for p in oracle.batch_files_sorted(str(tmp_path))]
    assert names == ["1.mtx", "2.mtx", "10.mtx"]


def test_mis_file_parse(oracle, tmp_path):
    p = tmp_path / "mis.txt"
    p.write_text("1\n3\n3\n9\n0\n")
    mis, card = oracle.read_mis(str(p), 5)
    assert mis.tolist() == [1, 0, 1, 0, 0] and card == 3
