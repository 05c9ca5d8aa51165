#!/usr/bin/env python3
"""Compaction of the device overlay into a new base on one MI355X.

Graph: R-MAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05), edgefactor 16,
symmetrised and deduplicated (default scale 24), greedy ascending-id MIS.
Updates: half deletions of existing edges, half insertions of non-edges,
distinct, 2.5% of the base entries per part (three parts by default),
applied in 200,000-update queued batches; after each part the overlay is
compacted into a new base (base order minus deletions, then insertions in
insertion order, core/flash_compactor.h:20-105), which also rebuilds the
sorted row copy that adjacency searches use once rows are no longer
ascending.  The first compaction of a graph allocates its outputs, the
second its sorted copy (the first had none to release); from the third on
every output reuses a buffer an earlier compaction released.

Reported per compaction: wall time of misfit_compact (device work plus two
small read-backs), and the algorithmic bytes of the merge, B_c = 8(n+1) +
4 m_old + 4 overlay slots + 8(n+1) + 4 m_new (old and new offsets and
entries, the overlay pool read once), against 8 TB/s.  The full check
(independence, maximality, count consistency) must pass at the end.  Prints
one JSON line.

    python misfit_amd/compact_bench.py [--scale 24] [--parts 3]
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import misfit_amd as M  # noqa: E402

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--batch", type=int, default=200_000)
    ap.add_argument("--parts", type=int, default=3, help="compactions, each after 2.5%% of the base in updates")
    args = ap.parse_args()
    t = time.time()
    ni, nl = M.gen_rmat(args.scale, 16, args.seed)
    mis0 = M.greedy_mis(ni, nl)
    n, m0 = len(ni) - 1, len(nl)
    part_len = int(0.025 * m0)
    nup = part_len * args.parts
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, nup, args.seed)
    print("R-MAT scale %d: n=%d m=%d, %d updates (%.1fs)" % (args.scale, n, m0, nup, time.time() - t),
          file=sys.stderr, flush=True)
    out = []
    with M.Engine(0) as e:
        name = e.device_info()[0]
        e.load_csr(ni, nl)
        ups = M.make_updates(u, v, e.flags(u, v))
        e.set_mis(mis0)
        del nl
        m_old = m0
        for rep in range(args.parts):
            part = ups[rep * part_len:(rep + 1) * part_len]
            tickets, st = [], None
            for i in range(0, len(part), args.batch):
                tickets.append(e.submit(part[i:i + args.batch]))
                if len(tickets) > 2:
                    st = e.wait(tickets.pop(0))
            for x in tickets:
                st = e.wait(x)
            e.sync()
            pool = int(st["overlay_entries"])  # overlay pool slots in use
            t = time.perf_counter()
            m_new = e.compact()
            e.sync()
            ms = (time.perf_counter() - t) * 1e3
            bc = 8 * (n + 1) + 4 * m_old + 4 * pool + 8 * (n + 1) + 4 * m_new
            rec = {"compaction": rep + 1, "base_entries_before": int(m_old), "overlay_slots": pool,
                   "entries_after": int(m_new),
                   "ms": round(ms, 2), "algorithmic_bytes": int(bc),
                   "achieved_gbs": round(bc / (ms / 1e3) / 1e9, 1),
                   "frac_of_8tbs": round(bc / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
            print("compaction %d: %s" % (rep + 1, rec), file=sys.stderr, flush=True)
            out.append(rec)
            m_old = m_new
        conf, unc, bad = e.check()
        if (conf, unc, bad) != (0, 0, 0):
            raise SystemExit("check failed after compaction: %s" % ((conf, unc, bad),))
    print(json.dumps({"workload": "R-MAT scale %d (edgefactor 16, 0.57/0.19/0.19/0.05, seed %d), overlay of 2.5%% of "
                                  "the base per compaction, %d compactions, %d-update queued batches"
                                  % (args.scale, args.seed, args.parts, args.batch),
                      "device": name, "n": n, "compactions": out, "final_check": "independent, maximal"}))


if __name__ == "__main__":
    main()
