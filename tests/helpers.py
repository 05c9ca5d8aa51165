"""Shared drivers: replay the same batches through the oracle and the engine."""
import numpy as np


def sample_inputs(oracle, sample_dir):
    import os
    g = oracle.Graph.load_egr(os.path.join(sample_dir, "sample.egr"))
    mis, card = oracle.read_mis(os.path.join(sample_dir, "sample_mis.txt"), g.n)
    batches = [g.parse_batch_file(p) for p in oracle.batch_files_sorted(os.path.join(sample_dir, "batches"))]
    return g, mis, batches


# The rule's per-batch counters, reported alike by the oracle and the engine:
# insertions with both ends in the MIS (each evicts), admissions on that
# branch, deletions with exactly one end in the MIS, admissions on that
# branch, and completion additions.
RULE_COUNTERS = ("active_ins", "evictions", "admit_ins", "active_del", "admit_del", "completion_adds")


def counter_mismatch(want_stats, got_stats):
    """The first rule counter on which two runs' per-batch stats differ ("" when
    none): the oracle's (or a reference run's) first, the engine's second."""
    if len(want_stats) != len(got_stats):
        return "%d batches against %d" % (len(want_stats), len(got_stats))
    for b, (w, g) in enumerate(zip(want_stats, got_stats)):
        for k in RULE_COUNTERS:
            if w[k] != g[k]:
                return "batch %d %s: %d expected, %d reported" % (b, k, w[k], g[k])
    return ""


def oracle_replay(oracle, nindex, nlist, mis0, batches, completion=True, literal=False):
    """Per-batch MIS (list of uint8 arrays), stats and the compacted CSR."""
    g = oracle.Graph.from_csr(nindex, nlist)
    mis = np.array(mis0, np.uint8)
    out, stats = [], []
    for (u, v, ins) in batches:
        st = g.process_batch(mis, u, v, ins, completion=completion, literal=literal)
        out.append(mis.copy())
        stats.append(st)
    return out, stats, g


def engine_replay(misfit_amd, eng, nindex, nlist, mis0, batches, completion=True, mode=None, check=True):
    eng.load_csr(nindex, nlist)
    eng.set_mis(mis0)
    eng.set_rule(completion)
    if mode is not None:
        eng.set_mode(mode)
    out, stats, checks = [], [], []
    for (u, v, ins) in batches:
        st = eng.process_batch(misfit_amd.make_updates(u, v, ins))
        out.append(eng.get_mis())
        stats.append(st)
        if check:
            checks.append(eng.check())
    return out, stats, checks


def random_graph(rng, n, p):
    """Symmetric simple graph as CSR with rows ascending."""
    a = rng.random((n, n)) < p
    a = np.triu(a, 1)
    a = a | a.T
    nindex = np.zeros(n + 1, np.int64)
    rows = [np.nonzero(a[i])[0].astype(np.int32) for i in range(n)]
    for i in range(n):
        nindex[i + 1] = nindex[i] + len(rows[i])
    nlist = np.concatenate(rows) if nindex[-1] else np.zeros(0, np.int32)
    return nindex, nlist


def random_batches(rng, oracle, nindex, nlist, nb, bs, general=False):
    """Batches of random pairs; flags fixed against the loaded graph unless
    general=True (then flags are random, pairs may repeat inside a batch)."""
    n = len(nindex) - 1
    g = oracle.Graph.from_csr(nindex, nlist)
    out = []
    for _ in range(nb):
        u = rng.integers(0, n, bs).astype(np.int32)
        v = rng.integers(0, n, bs).astype(np.int32)
        keep = u != v
        u, v = u[keep], v[keep]
        if general:
            # repeat some pairs inside the batch, random flags
            k = len(u) // 4
            if k:
                idx = rng.integers(0, len(u), k)
                u = np.concatenate([u, v[idx]])
                v = np.concatenate([v, u[idx]])
            ins = (rng.random(len(u)) < 0.5).astype(np.uint8)
        else:
            ins = g.flags(u, v)
        out.append((u, v, ins))
    return out


def existing_edge_batches(rng, nindex, nlist, nb, bs):
    """Deletion-heavy batches: random existing entries plus random pairs."""
    n = len(nindex) - 1
    src = np.repeat(np.arange(n, dtype=np.int32), np.diff(nindex))
    out = []
    for _ in range(nb):
        k = bs // 2
        pick = rng.integers(0, len(nlist), k)
        u1, v1 = src[pick], nlist[pick]
        u2 = rng.integers(0, n, bs - k).astype(np.int32)
        v2 = rng.integers(0, n, bs - k).astype(np.int32)
        u = np.concatenate([u1, u2])
        v = np.concatenate([v1, v2])
        keep = u != v
        out.append((u[keep], v[keep]))
    return out
