#!/usr/bin/env python3
"""Open-loop serving benchmark on one MI355X (the serving configuration).

Graph: R-MAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05), edgefactor 16,
symmetrised and deduplicated (default scale 26: 67,108,864 vertices, int64
offsets on the device), greedy ascending-id MIS.  Requests: update batches
(half deletions of existing edges, half insertions of non-edges, flags fixed
against the loaded graph) with sizes log-uniform in [100, 1,000,000], arriving
as a Poisson process at each offered rate (updates/s).  The server compacts
the graph when its overlay passes 5% of the base.  Per rate: achieved
updates/s, p50/p99 latency from arrival to result, p50/p99 queueing, and the
compactions that ran.  Prints one JSON line.

    python misfit_amd/serve_bench.py [--scale 26] [--rates 2e7,5e7,1e8] [--seconds 3]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import misfit_amd as M  # noqa: E402


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else 0.0


def schedule(rng, rate, seconds, lo=100, hi=1_000_000):
    """Arrival times (s) and sizes of the requests of one offered rate."""
    sizes, times = [], []
    mean = (hi - lo) / np.log(hi / lo)  # mean of a log-uniform size
    t = 0.0
    while t < seconds:
        t += rng.exponential(mean / rate)
        sizes.append(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        times.append(t)
    return np.array(times), np.array(sizes, np.int64)


def run_rate(srv, gid, ups, cursor, times, sizes):
    """Submit each request at its arrival time from a submitter thread; wait
    for all.  Returns per-request stats and the wall time."""
    reqs = [None] * len(times)
    starts = [0.0] * len(times)
    offs = np.concatenate([[cursor], cursor + np.cumsum(sizes)])
    t0 = time.perf_counter()

    def submitter():
        for i, (t, k) in enumerate(zip(times, sizes)):
            now = time.perf_counter() - t0
            if t > now:
                time.sleep(t - now)
            starts[i] = time.perf_counter()
            reqs[i] = srv.submit(gid, ups[offs[i]:offs[i + 1]])

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    out = []
    for i, r in enumerate(reqs):
        st = srv.wait(r)
        st["arrival_to_submit_ms"] = (starts[i] - t0 - times[i]) * 1e3
        out.append(st)
    wall = time.perf_counter() - t0
    return out, wall, int(offs[-1])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--rates", default="2e7,5e7,1e8", help="offered rates, updates/s")
    ap.add_argument("--seconds", type=float, default=3.0, help="arrival window per rate")
    ap.add_argument("--compact", type=float, default=0.05)
    args = ap.parse_args()
    rates = [float(x) for x in args.rates.split(",")]
    rng = np.random.default_rng(args.seed)
    plans = [schedule(rng, r, args.seconds) for r in rates]
    total = int(sum(int(s.sum()) for _, s in plans))
    t = time.time()
    print("generating R-MAT scale %d" % args.scale, file=sys.stderr, flush=True)
    ni, nl = M.gen_rmat(args.scale, 16, args.seed)
    print("  %d entries (%.1fs); greedy MIS" % (len(nl), time.time() - t), file=sys.stderr, flush=True)
    mis0 = M.greedy_mis(ni, nl)
    print("  %d updates (%.1fs)" % (total, time.time() - t), file=sys.stderr, flush=True)
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, total, args.seed)
    print("workload: n=%d m=%d, %d updates in %d requests, generated in %.1fs" % (
        len(ni) - 1, len(nl), total, sum(len(s) for _, s in plans), time.time() - t), file=sys.stderr, flush=True)
    with M.Engine(0) as eng:  # flags against the loaded graph
        eng.load_csr(ni, nl)
        ins = eng.flags(u, v)
        name = eng.device_info()[0]
    ups = M.make_updates(u, v, ins)
    results = []
    with M.Server(0, compact_fraction=args.compact, max_inflight=2) as srv:
        print("loading the graph into the server", file=sys.stderr, flush=True)
        gid = srv.add_graph(ni, nl, mis0)
        del nl
        cursor = 0
        for rate, (times, sizes) in zip(rates, plans):
            stats, wall, cursor = run_rate(srv, gid, ups, cursor, times, sizes)
            lat = [s["latency_ms"] for s in stats]
            q = [s["queue_ms"] for s in stats]
            n_up = int(sizes.sum())
            res = {
                "offered_updates_per_s": rate, "requests": len(stats), "updates": n_up,
                "achieved_updates_per_s": round(n_up / wall, 1),
                "latency_p50_ms": round(pct(lat, 50), 3), "latency_p99_ms": round(pct(lat, 99), 3),
                "queue_p50_ms": round(pct(q, 50), 3), "queue_p99_ms": round(pct(q, 99), 3),
                "device_ms_mean": round(float(np.mean([s["device_ms"] for s in stats])), 3),
                "compactions": int(stats[-1]["compactions_before"] - stats[0]["compactions_before"]),
                "rounds_mean": round(float(np.mean([s["batch"]["rounds"] for s in stats])), 1),
            }
            print("rate %.3g: %s" % (rate, res), file=sys.stderr, flush=True)
            results.append(res)
        conf, unc, bad = srv.check(gid)
        if (conf, unc, bad) != (0, 0, 0):
            raise SystemExit("serving: final check failed: %s" % ((conf, unc, bad),))
    print(json.dumps({
        "workload": "R-MAT scale %d (edgefactor 16, 0.57/0.19/0.19/0.05, symmetrised, deduplicated, seed %d), "
                    "greedy MIS; Poisson arrivals, batch sizes log-uniform in [100, 1e6], half deletions of "
                    "existing edges / half insertions of non-edges; compaction at %.0f%% overlay"
                    % (args.scale, args.seed, args.compact * 100),
        "device": name, "graph_entries": int(ni[-1]), "final_check": "independent, maximal",
        "rates": results}))


if __name__ == "__main__":
    main()
