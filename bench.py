#!/usr/bin/env python3
"""Edge updates/s of exact incremental-MIS maintenance on one MI355X.

Workload (R-MAT scale 22): 4,194,304 vertices, edgefactor 16, quadrant
probabilities (0.57, 0.19, 0.19, 0.05), symmetrised and deduplicated, seed 7;
greedy ascending-id initial MIS; a stream of batches of 10,000 updates (half
deletions of existing edges, half insertions of non-edges, distinct, seeded
shuffle), insertion flags fixed against the loaded graph.

A step is one batch: overlay update, the reference repair rule replayed
exactly in batch order, and the maximality completion, all on the device with
the graph resident in HBM.  Two drivers run the same stream from the same
start state:
  plain   one batch at a time, host-driven speculative rounds (the
          reference's synchronous engine pattern);
  served  the queued driver: each batch is one stream-ordered sequence
          (prologue kernels, a fixed number of grid-wide round kernels that
          turn into no-ops once the batch is resolved, and a one-workgroup
          finisher that runs any remaining rounds and the completion) with
          no host synchronisation inside it; batches are queued back to back
          with host staging overlapped (pipeline depth 2, the reference's
          async pattern generalised).
`value` is the served throughput.  Both drivers must end with the same MIS,
the full check (independence, maximality, count consistency) runs after
every plain batch, and a bounded sample is replayed by the CPU oracle (the
reference's host path: OpenMP 2-hop rows plus sequential rule replay) as the
CPU baseline and compared bit for bit with the GPU on the same sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edge updates/s (MIS bit-exact) + p50/p99 batch latency, served vs plain"
HBM_PEAK_GBS = 8000.0


def stream_copy_gbs(mib=2048, reps=20):
    """Copy bandwidth of the STREAM-style calibration program (read + write
    bytes per second), run as a child process before the engine starts."""
    import subprocess
    exe = os.path.join(ROOT, "misfit_amd", "bin", "stream_copy")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe, str(mib), str(reps)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        print("stream_copy failed: %s" % r.stderr.strip(), file=sys.stderr, flush=True)
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def committed_traffic(args):
    """HBM bytes per batch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by
    the copy-kernel calibration) for the default workload, from the committed
    summary; None for any other workload."""
    if (args.scale, args.batch, args.seed) != (22, 10000, 7):
        return None, None
    rel = os.path.join("profiles", "queued_batch_rmat22_traffic.md")
    try:
        with open(os.path.join(ROOT, rel)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_batch:"):
                    return int(line.split(":")[1]), rel
    except OSError:
        pass
    return None, None


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else 0.0


def build_workload(M, scale, batch, nbatches, seed):
    t = time.time()
    ni, nl = M.gen_rmat(scale, 16, seed)
    mis0 = M.greedy_mis(ni, nl)
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, batch * nbatches, seed)
    return ni, nl, mis0, u, v, time.time() - t


def run_plain(M, eng, ni, nl, mis0, ups, W, K, batch, measure_bytes=True):
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_PLAIN)
    lat, dev, nbytes, stats = [], [], [], []
    eng.sync()
    t_all = 0.0
    for b in range(W + K):
        chunk = ups[b * batch:(b + 1) * batch]
        t0 = time.perf_counter()
        st = eng.process_batch(chunk)
        dt = time.perf_counter() - t0
        conf, unc, bad = eng.check()
        if conf or unc or bad:
            raise SystemExit("plain: check failed after batch %d: %s" % (b, (conf, unc, bad)))
        print("plain batch %d: %.2f ms, rounds %d, evictions %d, PU %d, ADM %d" % (
            b, dt * 1e3, st["rounds"], st["evictions"], st["maybe_uncovered"], st["admissible"]),
            file=sys.stderr, flush=True)
        if b >= W:
            t_all += dt
            lat.append(dt * 1e3)
            dev.append(st["device_ms"])
            stats.append(st)
            if measure_bytes:
                nbytes.append(eng.batch_bytes(8)[0])
    return {"lat_ms": lat, "dev_ms": dev, "bytes": nbytes, "stats": stats, "wall_s": t_all,
            "mis": eng.get_mis()}


def run_served(M, eng, ni, nl, mis0, ups, W, K, batch, depth=2):
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_QUEUED)
    for b in range(W):
        eng.process_batch(ups[b * batch:(b + 1) * batch])
    eng.sync()
    starts, lat, dev, tickets, stats = {}, [], [], {}, []
    t0 = time.perf_counter()
    nxt = W
    for b in range(W, W + K):
        while nxt < W + K and nxt < b + depth:
            starts[nxt] = time.perf_counter()
            tickets[nxt] = eng.submit(ups[nxt * batch:(nxt + 1) * batch])
            nxt += 1
        st = eng.wait(tickets[b])
        lat.append((time.perf_counter() - starts[b]) * 1e3)
        dev.append(st["device_ms"])
        stats.append(st)
    eng.sync()
    wall = time.perf_counter() - t0
    conf, unc, bad = eng.check()
    return {"lat_ms": lat, "dev_ms": dev, "stats": stats, "wall_s": wall, "mis": eng.get_mis(),
            "check": (conf, unc, bad)}


def cpu_baseline(M, eng, ni, nl, mis0, ups, budget_s):
    """The oracle's literal engine (OpenMP rows + sequential replay) on the
    first N updates of the stream as one batch, N doubled until the time
    budget is used; the GPU replays the same sample and must agree."""
    from oracle import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g0 = O.Graph.from_csr(ni, nl)
    best = None
    n_up = 250
    spent = 0.0
    while n_up <= len(ups):
        s = ups[:n_up]
        g = O.Graph.from_csr(ni, nl)
        m = mis0.copy()
        t = time.perf_counter()
        g.process_batch(m, s["u"], s["v"], s["ins"], completion=True, literal=True, threads=threads)
        dt = time.perf_counter() - t
        spent += dt
        best = (n_up, dt, m)
        if spent + 2.5 * dt > budget_s:
            break
        n_up *= 2
    n_up, dt, m_cpu = best
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.process_batch(ups[:n_up])
    same = bool(np.array_equal(eng.get_mis(), m_cpu))
    del g0
    return {"value": n_up / dt, "unit": "edge updates/s", "cores": threads, "kind": "port",
            "sample": "first %d updates of the stream as one batch on the loaded graph, literal 2-hop rows "
                      "(OpenMP, %d threads) + sequential rule replay + completion; %.2f s" % (n_up, threads, dt),
            "bit_exact_vs_gpu": same}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=95)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--batch", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--served-only", action="store_true", help="skip the plain driver (profiling runs)")
    ap.add_argument("--no-stream-copy", action="store_true", help="skip the copy-bandwidth calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    copy = None if args.no_stream_copy else stream_copy_gbs()
    if copy:
        print("stream copy: %.1f GB/s" % copy["copy_gbs"], file=sys.stderr, flush=True)
    import misfit_amd as M

    W, K, B = args.warmup, args.steps, args.batch
    ni, nl, mis0, u, v, gen_s = build_workload(M, args.scale, B, W + K, args.seed)
    print("workload: n=%d m=%d generated in %.1fs" % (len(ni) - 1, len(nl), gen_s), file=sys.stderr, flush=True)
    with M.Engine(0) as eng:
        name, tot, _ = eng.device_info()
        eng.load_csr(ni, nl)
        ins = eng.flags(u, v)  # fixed against the loaded graph
        ups = M.make_updates(u, v, ins)
        plain = None
        if not args.served_only:
            plain = run_plain(M, eng, ni, nl, mis0, ups, W, K, B)
            print("plain: %.1f updates/s" % (K * B / plain["wall_s"]), file=sys.stderr, flush=True)
        served = run_served(M, eng, ni, nl, mis0, ups, W, K, B)
        print("served: %.1f updates/s" % (K * B / served["wall_s"]), file=sys.stderr, flush=True)
        if served["check"] != (0, 0, 0):
            raise SystemExit("served: final check failed: %s" % (served["check"],))
        same = plain is None or bool(np.array_equal(plain["mis"], served["mis"]))
        if not same:
            raise SystemExit("served and plain MIS differ")
        cpu = None
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(M, eng, ni, nl, mis0, ups, args.cpu_budget)
            print("cpu baseline: %.1f updates/s" % cpu["value"], file=sys.stderr, flush=True)
            if not cpu["bit_exact_vs_gpu"]:
                raise SystemExit("GPU differs from the oracle on the CPU-baseline sample")
    served_ups = K * B / served["wall_s"]
    roof = None
    if plain is not None and plain["bytes"]:
        tot_bytes = float(np.sum(plain["bytes"]))
        tot_s = float(np.sum(served["dev_ms"])) / 1e3
        ach = tot_bytes / tot_s / 1e9
        traffic, traffic_src = committed_traffic(args)
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "queued batch (prologue, grid rounds, finisher) timed per batch with HIP events",
                "algorithmic_bytes_per_batch": round(tot_bytes / K, 1),
                "algorithmic_bytes_per_update": round(tot_bytes / (K * B), 2),
                "kernel_ms_mean": round(float(np.mean(served["dev_ms"])), 4),
                "stream_copy_gbs": copy["copy_gbs"] if copy else None,
                "frac_of_stream_copy": round(ach / copy["copy_gbs"], 5) if copy else None}
    cfg = {
        "workload": "R-MAT scale 22 (4,194,304 vertices, edgefactor 16, a,b,c,d=0.57,0.19,0.19,0.05, "
                    "symmetrised, deduplicated, seed %d), greedy ascending-id MIS, %d-update batches "
                    "(half deletions of existing edges, half insertions of non-edges)" % (args.seed, B),
        "graph_entries": int(len(nl)),
        "batch_updates": B,
        "served_updates_per_s": round(served_ups, 1),
        "served_p50_ms": round(pct(served["lat_ms"], 50), 3),
        "served_p99_ms": round(pct(served["lat_ms"], 99), 3),
        "served_kernel_ms_mean": round(float(np.mean(served["dev_ms"])), 3),
        "mis_plain_equals_served": same,
        "check_after_every_plain_batch": "independent, maximal, counts consistent" if plain else None,
        "rounds_mean": round(float(np.mean([s["rounds"] for s in served["stats"]])), 1),
        "evictions_mean": round(float(np.mean([s["evictions"] for s in served["stats"]])), 1),
        "device": name,
    }
    if plain is not None:
        plain_ups = K * B / plain["wall_s"]
        cfg.update({
            "plain_updates_per_s": round(plain_ups, 1),
            "plain_p50_ms": round(pct(plain["lat_ms"], 50), 3),
            "plain_p99_ms": round(pct(plain["lat_ms"], 99), 3),
            "served_over_plain": round(served_ups / plain_ups, 3),
        })
    out = {
        "metric": METRIC, "value": round(served_ups, 1), "unit": "edge updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(served["wall_s"] * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generators in misfit_amd/csrc/gen.cpp)", "config": cfg,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
