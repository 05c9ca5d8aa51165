#!/usr/bin/env python3
"""Edge updates/s of exact incremental-MIS maintenance on one MI355X.

Workload (R-MAT scale 22): 4,194,304 vertices, edgefactor 16, quadrant
probabilities (0.57, 0.19, 0.19, 0.05), symmetrised and deduplicated, seed 7;
greedy ascending-id initial MIS; a stream of batches of 10,000 updates (half
deletions of existing edges, half insertions of non-edges, distinct, seeded
shuffle), insertion flags fixed against the loaded graph.

A step is one batch: overlay update, the reference repair rule replayed
exactly in batch order, and the maximality completion, all on the device with
the graph resident in HBM and the timed batches uploaded to device memory
before the timed region (misfit_stage_batch).  Two drivers run the same
stream from the same start state:
  plain   one batch at a time, host-driven speculative rounds (the
          reference's synchronous engine pattern);
  served  the queued driver: each batch is one stream-ordered sequence
          (prologue kernels, a fixed number of grid-wide round kernels that
          turn into no-ops once the batch is resolved, and a one-workgroup
          finisher that runs any remaining rounds and the completion) with
          no host synchronisation inside it; batches are queued back to back
          (pipeline depth 2, the reference's async pattern generalised).
`value` is the served throughput; `served_host_handover_updates_per_s` is
the same driver with every batch handed over from host memory at submission
(the PCIe-inclusive rate).  Both drivers must end with the same MIS,
the full check (independence, maximality, count consistency) runs after
every plain batch, and a bounded sample is replayed by the CPU oracle (the
reference's host path: OpenMP 2-hop rows plus sequential rule replay) as the
CPU baseline and compared bit for bit with the GPU on the same sample.

Other workloads (`--config`), each printed as its own JSON line of the same
shape, with the same checks:
  c1  Erdos-Renyi n=100,000, 800,000 undirected edges, seed 42; 10 batches
      of 2,000 uniform vertex pairs (toggles); CPU baseline on the whole
      stream.
  c3  2-D random geometric graph, n=2^24, average degree 32, seed 11 (int64
      offsets); batches of 50,000 pairs with endpoints drawn proportionally
      to (degree + 1); CPU baseline on as many whole batches as its budget
      allows.
  c5  64 graphs held on the device at once, n log-uniform in [2^14, 2^22],
      type uniform over {Erdos-Renyi degree 16, R-MAT edgefactor 16, RGG
      degree 32}; batches of 1,000-100,000 updates (log-uniform) with both
      endpoints in each graph's lowest-degree quarter, interleaved across
      graphs.  Served: the serving layer with every graph resident and all
      requests submitted at once; plain: the graphs one after another, one
      batch at a time.  `value` is the aggregate served throughput; per-graph
      p50/p99 latency is reported.
  compact  overlay compaction (core/flash_compactor.h:20-105) on the R-MAT
      scale 22 graph with an overlay of 5% of its entries (the first updates
      of the default stream): the GPU merges it into a new base in HBM
      (value: ms per compaction) and the reference's own compact_flash_graph,
      built from its sources into oracle/_ref, compacts the same base and
      overlay on the host (the CPU baseline, kind "reference"); the two
      compacted graphs must be identical.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


class _Ranges:
    """Named ranges for rocprofv3 --marker-trace around the timed regions
    (roctx from the ROCm install; nothing happens when it is absent)."""

    def __init__(self):
        self.lib = None
        try:
            import ctypes
            root = os.environ.get("ROCM_PATH", "/opt/rocm")
            self.lib = ctypes.CDLL(os.path.join(root, "lib", "librocprofiler-sdk-roctx.so.1"))
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        except OSError:
            self.lib = None

    def push(self, name):
        if self.lib:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib:
            self.lib.roctxRangePop()


RANGES = _Ranges()

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


def committed_figure(rel, key):
    """The integer after `key:` at the start of a line of a committed rocprofv3
    summary, with the summary's path; (None, None) when either is missing."""
    try:
        with open(os.path.join(ROOT, rel)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":", 1)[1]), rel
    except OSError:
        pass
    return None, None


def committed_traffic(args):
    """HBM bytes per batch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by
    the copy-kernel calibration) for the default workload, from the committed
    summary; None for any other workload."""
    if (args.scale, args.batch, args.seed) != (22, 10000, 7):
        return None, None
    return committed_figure(os.path.join("profiles", "queued_batch_rmat22_traffic.md"), "traffic_bytes_per_batch")


KERNEL_NOTE = ("queued batch (prologue, grid rounds, finisher): device time of the timed region by a pair of HIP "
               "events on the engine's stream, divided by the batches in it")


def physical_cores():
    """Physical cores among the CPUs this process may run on: distinct
    (package, core) pairs in /proc/cpuinfo (the logical CPU count where that
    file does not say)."""
    allowed = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    cores, proc, phys = set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "processor":
                    proc, phys = int(val), None
                elif key == "physical id":
                    phys = val
                elif key == "core id" and proc is not None and (allowed is None or proc in allowed):
                    cores.add((phys, val))
    except (OSError, ValueError):
        pass
    return len(cores) or (len(allowed) if allowed else os.cpu_count() or 1)


def cpu_threads():
    """Threads of the CPU baseline: one per physical core (at most
    OMP_NUM_THREADS when that is set)."""
    n = physical_cores()
    try:
        return max(1, min(n, int(os.environ["OMP_NUM_THREADS"])))
    except (KeyError, ValueError):
        return n


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs, physical cores."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "physical_cores": physical_cores()}


def phase_means(stats):
    """Mean device-clock split of served batches: the prologue (overlay
    update, may-uncover set, admissible index), the speculative rounds, the
    rule counters with the completion's candidates, and the completion with
    the publication."""
    if not stats:
        return None
    pro = float(np.mean([s["prologue_ms"] for s in stats]))
    rnd = float(np.mean([s["rounds_ms"] for s in stats]))
    cnt = float(np.mean([s["counters_ms"] for s in stats]))
    dev = float(np.mean([s["device_ms"] for s in stats]))
    return {"prologue": round(pro, 4), "rounds": round(rnd, 4), "counters": round(cnt, 4),
            "completion_and_publication": round(dev - pro - rnd - cnt, 4)}


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else 0.0


def build_workload(M, scale, batch, nbatches, seed):
    t = time.time()
    ni, nl = M.gen_rmat(scale, 16, seed)
    mis0 = M.greedy_mis(ni, nl)
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, batch * nbatches, seed)
    return ni, nl, mis0, u, v, time.time() - t


def stage_all(eng, ups, nb, batch):
    """Uploads every batch of the stream to device memory ahead of the timed
    region, so that the timed batches start from inputs resident in HBM."""
    return [eng.stage(ups[b * batch:(b + 1) * batch]) for b in range(nb)]


def release_all(eng, hs):
    for h in hs:
        eng.release_staged(h)


def run_plain(M, eng, ni, nl, mis0, ups, W, K, batch, measure_bytes=True, index_bytes=8):
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_PLAIN)
    lat, dev, nbytes, stats, all_bytes = [], [], [], [], []
    hs = stage_all(eng, ups, W + K, batch)
    eng.sync()
    t_all = 0.0
    for b in range(W + K):
        t0 = time.perf_counter()
        st = eng.wait(eng.submit_staged(hs[b]))
        dt = time.perf_counter() - t0
        conf, unc, bad = eng.check()
        if conf or unc or bad:
            raise SystemExit("plain: check failed after batch %d: %s" % (b, (conf, unc, bad)))
        print("plain batch %d: %.2f ms, rounds %d, evictions %d, PU %d, ADM %d" % (
            b, dt * 1e3, st["rounds"], st["evictions"], st["maybe_uncovered"], st["admissible"]),
            file=sys.stderr, flush=True)
        nb = eng.batch_bytes(index_bytes)[0] if measure_bytes else None
        all_bytes.append(nb)
        if b >= W:
            t_all += dt
            lat.append(dt * 1e3)
            dev.append(st["device_ms"])
            stats.append(st)
            if measure_bytes:
                nbytes.append(nb)
    release_all(eng, hs)
    return {"lat_ms": lat, "dev_ms": dev, "bytes": nbytes, "bytes_all": all_bytes, "stats": stats,
            "wall_s": t_all, "mis": eng.get_mis()}


def run_served(M, eng, ni, nl, mis0, ups, W, K, batch, depth=2, staged=True):
    """Queued mode, depth batches in flight.  staged: the batches are in
    device memory before the timed region (submit_staged); otherwise each is
    handed over from host memory by submit (the PCIe-inclusive rate)."""
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_QUEUED)
    for b in range(W):
        eng.process_batch(ups[b * batch:(b + 1) * batch])
    hs = [None] * W + [eng.stage(ups[b * batch:(b + 1) * batch]) for b in range(W, W + K)] if staged else None
    eng.sync()
    starts, lat, dev, tickets, stats, sub = {}, [], [], {}, [], []
    eng.timer_begin()  # HIP event on the engine's stream: the timed region's device time
    RANGES.push("bench:served")
    t0 = time.perf_counter()
    nxt = W
    for b in range(W, W + K):
        while nxt < W + K and nxt < b + depth:
            starts[nxt] = time.perf_counter()
            tickets[nxt] = (eng.submit_staged(hs[nxt]) if staged
                            else eng.submit(ups[nxt * batch:(nxt + 1) * batch]))
            sub.append((time.perf_counter() - starts[nxt]) * 1e3)
            nxt += 1
        st = eng.wait(tickets[b])
        lat.append((time.perf_counter() - starts[b]) * 1e3)
        dev.append(st["device_ms"])
        stats.append(st)
    eng.sync()
    wall = time.perf_counter() - t0
    RANGES.pop()
    region_ms = eng.timer_end()
    if staged:
        release_all(eng, hs[W:])
    conf, unc, bad = eng.check()
    return {"lat_ms": lat, "dev_ms": dev, "stats": stats, "wall_s": wall, "mis": eng.get_mis(),
            "check": (conf, unc, bad), "region_ms": region_ms, "submit_ms": sub}


def cpu_baseline(M, eng, ni, nl, mis0, ups, budget_s):
    """The oracle's literal engine (OpenMP rows + sequential replay) on the
    first N updates of the stream as one batch, N doubled until the time
    budget is used; the GPU replays the same sample and must agree."""
    from oracle import pyoracle as O
    threads = cpu_threads()
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
    return {"value": n_up / dt, "unit": "edge updates/s", "cores": threads, "kind": "port", **host_info(),
            "sample": "first %d updates of the stream as one batch on the loaded graph, literal 2-hop rows "
                      "(OpenMP, %d threads) + sequential rule replay + completion; %.2f s" % (n_up, threads, dt),
            "bit_exact_vs_gpu": same}


def open_loop(M, eng, ni, nl, mis0, ups, batch, nb, rate, mode, depth=2, seed=1):
    """Batches arriving as a Poisson process at `rate` batches/s, from the
    initial state; latency is arrival -> result (host queueing included).
    plain: a batch starts when it has arrived and the previous one is done;
    served: queued mode, a batch is submitted on arrival while fewer than
    `depth` are on the device, completions are polled without blocking."""
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate, nb))
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_PLAIN if mode == "plain" else M.MODE_QUEUED)
    eng.sync()
    lat = []
    t0 = time.perf_counter()
    if mode == "plain":
        for b in range(nb):
            while time.perf_counter() - t0 < arr[b]:
                pass
            eng.process_batch(ups[b * batch:(b + 1) * batch])
            lat.append((time.perf_counter() - t0 - arr[b]) * 1e3)
    else:
        tickets, nxt, done = [], 0, 0
        while done < nb:
            now = time.perf_counter() - t0
            if nxt < nb and arr[nxt] <= now and nxt - done < depth:
                tickets.append(eng.submit(ups[nxt * batch:(nxt + 1) * batch]))
                nxt += 1
                continue
            if done < nxt and eng.poll(tickets[done]):
                eng.wait(tickets[done])
                lat.append((time.perf_counter() - t0 - arr[done]) * 1e3)
                done += 1
    eng.sync()
    return {"p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3), "mis": eng.get_mis()}


def latency_at_load(M, eng, ni, nl, mis0, ups, batch, nb, plain_ups, fracs=(0.5, 0.8, 1.2)):
    """Open-loop p50/p99 of both drivers at the same offered loads (fractions
    of the plain driver's measured throughput; above 1 the plain driver's
    queue grows for as long as batches arrive); both must end with the same
    MIS."""
    out = []
    for f in fracs:
        rate = f * plain_ups / batch
        p = open_loop(M, eng, ni, nl, mis0, ups, batch, nb, rate, "plain")
        s = open_loop(M, eng, ni, nl, mis0, ups, batch, nb, rate, "served")
        if not np.array_equal(p["mis"], s["mis"]):
            raise SystemExit("open loop: served and plain MIS differ")
        out.append({"offered_updates_per_s": round(rate * batch, 1), "fraction_of_plain": f, "batches": nb,
                    "plain_p50_ms": p["p50_ms"], "plain_p99_ms": p["p99_ms"],
                    "served_p50_ms": s["p50_ms"], "served_p99_ms": s["p99_ms"]})
        print("open loop at %.3g updates/s: plain p50/p99 %.3f/%.3f ms, served %.3f/%.3f ms" % (
            rate * batch, p["p50_ms"], p["p99_ms"], s["p50_ms"], s["p99_ms"]), file=sys.stderr, flush=True)
    return out


def workload_c1(M, W, K):
    t = time.time()
    ni, nl = M.gen_er(100_000, 800_000, 42)
    mis0 = M.greedy_mis(ni, nl)
    u, v = M.gen_updates(M.UNIFORM_PAIRS, ni, nl, 2000 * (W + K), 42)
    return ni, nl, mis0, u, v, time.time() - t


def workload_c3(M, W, K):
    t = time.time()
    ni, nl = M.gen_rgg(1 << 24, 32.0, 11)
    mis0 = M.greedy_mis(ni, nl)
    u, v = M.gen_updates(M.PREFERENTIAL, ni, nl, 50_000 * (W + K), 11)
    return ni, nl, mis0, u, v, time.time() - t


def cpu_stream(M, eng, ni, nl, mis0, ups, batch, nb, budget_s, batch_bytes=None):
    """The oracle's literal engine replaying whole batches of the stream in
    order from the initial MIS until the time budget is used (at least one
    batch); the GPU replays the same batches and must agree bit for bit.
    batch_bytes: the algorithmic bytes of each batch (for the bytes rate)."""
    from oracle import pyoracle as O
    threads = cpu_threads()
    g = O.Graph.from_csr(ni, nl)
    m = mis0.copy()
    lat = []
    for b in range(nb):
        s = ups[b * batch:(b + 1) * batch]
        t = time.perf_counter()
        g.process_batch(m, s["u"], s["v"], s["ins"], completion=True, literal=True, threads=threads)
        lat.append(time.perf_counter() - t)
        if sum(lat) + 1.5 * lat[-1] > budget_s:
            break
    done = len(lat)
    eng.load_csr(ni, nl)
    eng.set_mis(mis0)
    eng.set_mode(M.MODE_PLAIN)
    for b in range(done):
        eng.process_batch(ups[b * batch:(b + 1) * batch])
    same = bool(np.array_equal(eng.get_mis(), m))
    del g
    out = {"value": done * batch / sum(lat), "unit": "edge updates/s", "cores": threads, "kind": "port",
           **host_info(),
           "sample": "first %d of %d batches of the stream, whole and in order, literal 2-hop rows (OpenMP, %d "
                     "threads) + sequential rule replay + completion; %.2f s" % (done, nb, threads, sum(lat)),
           "batches": done, "per_batch_ms": [round(x * 1e3, 2) for x in lat],
           "p50_ms": round(pct([x * 1e3 for x in lat], 50), 3), "p99_ms": round(pct([x * 1e3 for x in lat], 99), 3),
           "bit_exact_vs_gpu": same}
    if batch_bytes is not None and len(batch_bytes) >= done:
        out["algorithmic_bytes_per_s"] = round(float(np.sum(batch_bytes[:done])) / sum(lat), 1)
    return out


def single_graph_line(M, args, name, ni, nl, mis0, u, v, W, K, B, index_bytes, copy, workload, cpu_mode):
    """Plain and served runs of one graph's stream, the checks, the CPU
    baseline and the JSON line."""
    with M.Engine(0) as eng:
        dev_name, _, _ = eng.device_info()
        eng.load_csr(ni, nl)
        ins = eng.flags(u, v)  # fixed against the loaded graph
        ups = M.make_updates(u, v, ins)
        if args.served_only:  # profiling runs: the served batches alone, with their final check
            served = run_served(M, eng, ni, nl, mis0, ups, W, K, B)
            if served["check"] != (0, 0, 0):
                raise SystemExit("served: final check failed: %s" % (served["check"],))
            return {"metric": "edge updates/s (served batches only: a profiling run)",
                    "value": round(K * B / served["wall_s"], 1), "unit": "edge updates/s", "n_gpus": 1,
                    "steps": K, "warmup": W, "config": {"workload": workload, "name": name, "device": dev_name}}
        plain = run_plain(M, eng, ni, nl, mis0, ups, W, K, B, index_bytes=index_bytes)
        print("plain: %.1f updates/s" % (K * B / plain["wall_s"]), file=sys.stderr, flush=True)
        served = run_served(M, eng, ni, nl, mis0, ups, W, K, B)
        print("served: %.1f updates/s" % (K * B / served["wall_s"]), file=sys.stderr, flush=True)
        if served["check"] != (0, 0, 0):
            raise SystemExit("served: final check failed: %s" % (served["check"],))
        if not np.array_equal(plain["mis"], served["mis"]):
            raise SystemExit("served and plain MIS differ")
        served_h = run_served(M, eng, ni, nl, mis0, ups, W, K, B, staged=False)
        print("served, host hand-over: %.1f updates/s" % (K * B / served_h["wall_s"]), file=sys.stderr, flush=True)
        if served_h["check"] != (0, 0, 0) or not np.array_equal(plain["mis"], served_h["mis"]):
            raise SystemExit("served with host hand-over: final check failed or MIS differs")
        plain_ups0 = K * B / plain["wall_s"]
        loads = latency_at_load(M, eng, ni, nl, mis0, ups, B, W + K, plain_ups0)
        cpu = None
        if not args.no_cpu_baseline:
            if cpu_mode == "stream":
                cpu = cpu_stream(M, eng, ni, nl, mis0, ups, B, W + K, args.cpu_budget, plain["bytes_all"])
            else:
                cpu = cpu_baseline(M, eng, ni, nl, mis0, ups, args.cpu_budget)
            print("cpu baseline: %.1f updates/s" % cpu["value"], file=sys.stderr, flush=True)
            if not cpu["bit_exact_vs_gpu"]:
                raise SystemExit("GPU differs from the oracle on the CPU-baseline sample")
    served_ups = K * B / served["wall_s"]
    plain_ups = K * B / plain["wall_s"]
    tot_bytes = float(np.sum(plain["bytes"]))
    ach = tot_bytes / (served["region_ms"] / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "traffic_source": None,
            "kernel": KERNEL_NOTE,
            "algorithmic_bytes_per_batch": round(tot_bytes / K, 1),
            "algorithmic_bytes_per_update": round(tot_bytes / (K * B), 2),
            "kernel_ms_mean": round(served["region_ms"] / K, 4),
            "batch_device_clock_ms_mean": round(float(np.mean(served["dev_ms"])), 4),
            "stream_copy_gbs": copy["copy_gbs"] if copy else None,
            "frac_of_stream_copy": round(ach / copy["copy_gbs"], 5) if copy else None}
    cfg = {"workload": workload, "name": name, "graph_entries": int(len(nl)), "batch_updates": B,
           "served_updates_per_s": round(served_ups, 1),
           "served_p50_ms": round(pct(served["lat_ms"], 50), 3), "served_p99_ms": round(pct(served["lat_ms"], 99), 3),
           "served_kernel_ms_mean": round(float(np.mean(served["dev_ms"])), 3),
           "served_phase_ms_mean": phase_means(served["stats"]),
           "served_submit_ms_mean": round(float(np.mean(served["submit_ms"])), 4),
           "served_host_handover_updates_per_s": round(K * B / served_h["wall_s"], 1),
           "plain_updates_per_s": round(plain_ups, 1),
           "plain_p50_ms": round(pct(plain["lat_ms"], 50), 3), "plain_p99_ms": round(pct(plain["lat_ms"], 99), 3),
           "served_over_plain": round(served_ups / plain_ups, 3), "mis_plain_equals_served": True,
           "latency_note": "served_p50/p99: closed loop, %d batches in flight, submit -> result including "
                           "the wait behind the batch ahead; open_loop: both drivers at the same offered "
                           "load, arrival -> result" % 2,
           "open_loop": loads,
           "check_after_every_plain_batch": "independent, maximal, counts consistent",
           "rounds_mean": round(float(np.mean([s["rounds"] for s in served["stats"]])), 1),
           "device": dev_name}
    return {"metric": METRIC, "value": round(served_ups, 1), "unit": "edge updates/s", "n_gpus": 1,
            "steps": K, "warmup": W, "ms_per_step": round(served["wall_s"] * 1e3 / K, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (seeded generators in misfit_amd/csrc/gen.cpp)", "config": cfg,
            "roofline": roof, "cpu_baseline": cpu}


def compact_line(M, args, copy):
    """GPU compaction of a 5% overlay on the R-MAT scale 22 graph, timed over
    `steps` fresh loads of the same base and overlay; the reference's
    compactor on the same input as the CPU baseline; identical results."""
    import shutil
    import tempfile
    ni, nl = M.gen_rmat(22, 16, args.seed)
    mis0 = M.greedy_mis(ni, nl)
    n, m = len(ni) - 1, len(nl)
    nup = int(0.05 * m)
    u, v = M.gen_updates(M.MIXED_EXISTING, ni, nl, nup, args.seed)
    print("workload: n=%d m=%d, %d updates" % (n, m, nup), file=sys.stderr, flush=True)
    ms, pool, new_m, dt = [], 0, 0, 0.0
    tmpdir = tempfile.mkdtemp(dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        with M.Engine(0) as eng:
            name = eng.device_info()[0]
            eng.load_csr(ni, nl)
            ins = eng.flags(u, v)
            ups = M.make_updates(u, v, ins)
            for step in range(args.warmup + args.steps):
                eng.load_csr(ni, nl)
                eng.set_mis(mis0)
                tickets, st = [], None
                for i in range(0, len(ups), 100_000):
                    tickets.append(eng.submit(ups[i:i + 100_000]))
                    if len(tickets) > 2:
                        st = eng.wait(tickets.pop(0))
                for t in tickets:
                    st = eng.wait(t)
                eng.sync()
                pool = int(st["overlay_entries"])
                t0 = time.perf_counter()
                new_m = eng.compact()
                eng.sync()
                dt = (time.perf_counter() - t0) * 1e3
                if step >= args.warmup:
                    ms.append(dt)
                print("compaction %d: %.2f ms" % (step, dt), file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            cni, cnl = eng.export_csr()
            M.write_egr(os.path.join(tmpdir, "gpu.egr"), cni, cnl)
            with_write = dt + (time.perf_counter() - t0) * 1e3
            conf, unc, bad = eng.check()
            if (conf, unc, bad) != (0, 0, 0):
                raise SystemExit("check failed after compaction: %s" % ((conf, unc, bad),))
        cpu = None
        if not args.no_cpu_baseline:
            from oracle import pyoracle as O
            if not O.ref_available():
                raise SystemExit("oracle/_ref (the reference's compactor) is not built")
            ref_path = os.path.join(tmpdir, "ref.egr")
            M.write_egr(ref_path, ni, nl)
            ref = O.RefGraph(ref_path)
            # every update of this stream changes the graph (deletions of edges,
            # insertions of non-edges, each pair once), so the reference's
            # row-scanning adjacency tests are skipped while its overlay is built
            ref.apply(u, v, ins, check=False)
            t0 = time.perf_counter()
            ref.compact_to(ref_path)
            ref_ms = (time.perf_counter() - t0) * 1e3
            ref.close()
            rni, rnl, _ = M.read_egr(ref_path)
            same = bool(np.array_equal(rni, cni) and np.array_equal(rnl, cnl))
            if not same:
                raise SystemExit("GPU compaction differs from the reference's")
            cpu = {"value": round(ref_ms, 2), "unit": "ms", "cores": 1, "kind": "reference",
                   "sample": "the reference's compact_flash_graph (core/flash_compactor.h, built from its sources "
                             "into oracle/_ref) on the same base and overlay, writing the wide .egr to %s"
                             % ("/dev/shm" if tmpdir.startswith("/dev/shm") else "a temporary directory"),
                   "identical_to_gpu": same}
            print("reference compaction: %.1f ms" % ref_ms, file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    mean = float(np.mean(ms))
    bc = 8 * (n + 1) + 4 * m + 4 * pool + 8 * (n + 1) + 4 * int(new_m)
    # HBM bytes of one compaction (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated), committed
    # for this workload
    traffic, traffic_src = None, None
    if args.seed == 7:
        traffic, traffic_src = committed_figure(os.path.join("profiles", "compaction_rmat22_traffic.md"),
                                                "traffic_bytes_per_compaction")
    ach = bc / (mean / 1e3) / 1e9
    return {"metric": "overlay compaction time (merge into a new base, sorted row copy rebuilt)",
            "value": round(mean, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(mean, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic (seeded generators in misfit_amd/csrc/gen.cpp)",
            "config": {"workload": "R-MAT scale 22 (seed %d), greedy MIS, overlay from the first %d updates of the "
                                   "default stream (5%% of the entries), applied in 100,000-update queued batches"
                                   % (args.seed, nup),
                       "name": "compact", "graph_entries": m, "overlay_slots": pool, "entries_after": int(new_m),
                       "gpu_compact_plus_export_and_write_ms": round(with_write, 2), "device": name},
            "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                         "kernel": "misfit_compact (all its kernels) timed by the host around the call",
                         "algorithmic_bytes_per_compaction": int(bc),
                         "stream_copy_gbs": copy["copy_gbs"] if copy else None},
            "cpu_baseline": cpu}


C5_TYPES = ("er", "rmat", "rgg")


def workload_c5(M, n_graphs, per_graph, seed=5):
    """Graph sizes, types and batch sizes of the many-graph workload."""
    rng = np.random.default_rng(seed)
    specs = []
    for g in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(2 ** 14), np.log(2 ** 22))))
        kind = C5_TYPES[int(rng.integers(0, 3))]
        sizes = [int(np.exp(rng.uniform(np.log(1000), np.log(100_000)))) for _ in range(per_graph)]
        specs.append((kind, n, sizes, seed * 1000 + g))
    return specs


def gen_graph_c5(M, kind, n, gseed):
    if kind == "er":
        return M.gen_er(n, 8 * n, gseed)
    if kind == "rmat":
        return M.gen_rmat(int(round(np.log2(n))), 16, gseed)
    return M.gen_rgg(n, 32.0, gseed)


def many_graph_line(M, args, copy):
    """64 graphs: plain (one graph after another, one batch at a time, check
    after every batch) and served (all graphs resident, every request
    submitted at once, interleaved across graphs)."""
    specs = workload_c5(M, args.graphs, args.steps)
    graphs = []
    t0 = time.time()
    with M.Engine(0) as eng:
        dev_name, _, _ = eng.device_info()
        for i, (kind, n, sizes, gseed) in enumerate(specs):
            ni, nl = gen_graph_c5(M, kind, n, gseed)
            mis0 = M.greedy_mis(ni, nl)
            u, v = M.gen_updates(M.LOW_DEGREE, ni, nl, int(sum(sizes)), gseed)
            eng.load_csr(ni, nl)
            ins = eng.flags(u, v)
            ups = M.make_updates(u, v, ins)
            offs = np.concatenate([[0], np.cumsum(sizes)])
            graphs.append({"kind": kind, "n": len(ni) - 1, "ni": ni, "nl": nl, "mis0": mis0,
                           "batches": [ups[offs[b]:offs[b + 1]] for b in range(len(sizes))]})
            print("graph %d: %s n=%d m=%d, %d updates (%.1fs)" % (i, kind, len(ni) - 1, len(nl), len(ups),
                                                                 time.time() - t0), file=sys.stderr, flush=True)
        total_up = sum(len(b) for g in graphs for b in g["batches"])
        # plain: one graph after another (loads are not timed)
        plain_s, plain_lat, tot_bytes = 0.0, [], 0.0
        for i, g in enumerate(graphs):
            eng.load_csr(g["ni"], g["nl"])
            eng.set_mis(g["mis0"])
            eng.set_mode(M.MODE_PLAIN)
            eng.sync()
            for b in g["batches"]:
                t = time.perf_counter()
                eng.process_batch(b)
                dt = time.perf_counter() - t
                plain_s += dt
                plain_lat.append(dt * 1e3)
                tot_bytes += eng.batch_bytes(8)[0]
                conf, unc, bad = eng.check()
                if conf or unc or bad:
                    raise SystemExit("plain: check failed on graph %d: %s" % (i, (conf, unc, bad)))
            g["mis_plain"] = eng.get_mis()
            print("plain graph %d: %d batches, %.1f ms" % (i, len(g["batches"]), sum(plain_lat[-len(g["batches"]):])),
                  file=sys.stderr, flush=True)
        print("plain: %.1f updates/s" % (total_up / plain_s), file=sys.stderr, flush=True)
    def serve_closed(persistent):
        """Every request submitted at once, interleaved across graphs."""
        with M.Server(0, compact_fraction=0.05, max_inflight=args.inflight, workers=args.workers,
                      persistent=persistent) as srv:
            gids = [srv.add_graph(g["ni"], g["nl"], g["mis0"]) for g in graphs]
            order = [(gi, b) for b in range(args.steps) for gi in range(len(graphs))]
            RANGES.push("bench:served_c5")
            t = time.perf_counter()
            reqs = [(gi, srv.submit(gids[gi], graphs[gi]["batches"][b])) for gi, b in order]
            st = [(gi, srv.wait(r)) for gi, r in reqs]
            w = time.perf_counter() - t
            RANGES.pop()
            for gi, g in enumerate(graphs):
                if not np.array_equal(srv.get_mis(gids[gi], g["n"]), g["mis_plain"]):
                    raise SystemExit("served (persistent %d) and plain MIS differ on graph %d" % (persistent, gi))
                if srv.check(gids[gi]) != (0, 0, 0):
                    raise SystemExit("served (persistent %d): final check failed on graph %d" % (persistent, gi))
        return st, w

    def serve_open(persistent, rate, seed):
        """Open loop: each graph's requests arrive as a Poisson process, the
        graphs' rates in proportion to their updates and `rate` updates/s in
        all; latency from arrival to result."""
        rng = np.random.default_rng(seed)
        arrivals = []
        for gi, g in enumerate(graphs):
            sizes = [len(b) for b in g["batches"]]
            share = rate * sum(sizes) / total_up  # this graph's updates/s
            t = 0.0
            for b, k in enumerate(sizes):
                t += rng.exponential(k / share)
                arrivals.append((t, gi, b))
        arrivals.sort()
        with M.Server(0, compact_fraction=0.05, max_inflight=args.inflight, workers=args.workers,
                      persistent=persistent) as srv:
            gids = [srv.add_graph(g["ni"], g["nl"], g["mis0"]) for g in graphs]
            lat = {gi: [] for gi in range(len(graphs))}
            reqs = []
            t0 = time.perf_counter()
            for t, gi, b in arrivals:
                now = time.perf_counter() - t0
                if t > now:
                    time.sleep(t - now)
                sub = time.perf_counter() - t0
                reqs.append((gi, sub - t, srv.submit(gids[gi], graphs[gi]["batches"][b])))
            for gi, late, r in reqs:
                st = srv.wait(r)
                lat[gi].append(st["latency_ms"] + 1e3 * late)  # from the arrival, not the (later) submission
            span = time.perf_counter() - t0
            for gi, g in enumerate(graphs):
                if not np.array_equal(srv.get_mis(gids[gi], g["n"]), g["mis_plain"]):
                    raise SystemExit("open loop: served and plain MIS differ on graph %d" % gi)
        allv = [x for v in lat.values() for x in v]
        per = [(gi, round(pct(v, 50), 3), round(pct(v, 99), 3)) for gi, v in lat.items()]
        return {"offered_updates_per_s": round(rate, 1), "last_arrival_s": round(arrivals[-1][0], 4),
                "span_s": round(span, 4),
                "latency_p50_ms": round(pct(allv, 50), 3), "latency_p99_ms": round(pct(allv, 99), 3),
                "per_graph_p50_p99_ms": per}

    # the first server pass in a process pays one-off loading costs: a
    # warm-up pass with the same settings goes first
    serve_closed(args.persistent)
    stats, wall = serve_closed(args.persistent)
    # device time with at least one batch in progress, over the served span
    # (batch intervals from the device clock; the serving kernel itself is
    # resident the whole time, so a kernel trace shows the GPU always busy)
    iv = sorted((st["batch"]["device_start_ms"], st["batch"]["device_end_ms"]) for _, st in stats)
    busy, cur_s, cur_e = 0.0, None, None
    for a, b in iv:
        if cur_e is None or a > cur_e:
            if cur_e is not None:
                busy += cur_e - cur_s
            cur_s, cur_e = a, b
        else:
            cur_e = max(cur_e, b)
    if cur_e is not None:
        busy += cur_e - cur_s
    dev_span = max(b for _, b in iv) - iv[0][0] if iv else 0.0  # first start to the latest end
    busy_frac = busy / dev_span if dev_span > 0 else 0.0
    concurrency = sum(b - a for a, b in iv) / dev_span if dev_span > 0 else 0.0
    _, wall_streams = serve_closed(0)
    print("served: %.1f updates/s (serving kernel), %.1f (graphs' streams only)"
          % (total_up / wall, total_up / wall_streams), file=sys.stderr, flush=True)
    open_loop = []
    for frac in (0.5, 0.8):
        ol = serve_open(args.persistent, frac * total_up / wall, 11 + int(10 * frac))
        ol["fraction_of_served"] = frac
        open_loop.append(ol)
        print("open loop at %.3g updates/s: p50/p99 %.3f/%.3f ms" % (ol["offered_updates_per_s"], ol["latency_p50_ms"],
                                                                    ol["latency_p99_ms"]), file=sys.stderr, flush=True)
    served_ups = total_up / wall
    per = []
    for gi, g in enumerate(graphs):
        lat = [s["latency_ms"] for k, s in stats if k == gi]
        dev = [s["device_ms"] for k, s in stats if k == gi]
        per.append({"graph": gi, "type": g["kind"], "n": g["n"], "entries": int(len(g["nl"])),
                    "updates": int(sum(len(b) for b in g["batches"])),
                    "latency_p50_ms": round(pct(lat, 50), 3), "latency_p99_ms": round(pct(lat, 99), 3),
                    "device_p50_ms": round(pct(dev, 50), 3), "device_p99_ms": round(pct(dev, 99), 3)})
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_many(M, graphs, args.cpu_budget)
        print("cpu baseline: %.1f updates/s" % cpu["value"], file=sys.stderr, flush=True)
        if not cpu["bit_exact_vs_gpu"]:
            raise SystemExit("GPU differs from the oracle on the CPU-baseline sample")
    ach = tot_bytes / wall / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "traffic_source": None,
            "kernel": "all graphs' batches, served concurrently: algorithmic bytes over the served wall time",
            "algorithmic_bytes_per_update": round(tot_bytes / total_up, 2),
            "stream_copy_gbs": copy["copy_gbs"] if copy else None}
    lat_all = [s["latency_ms"] for _, s in stats]
    cfg = {"workload": "%d graphs on the device at once, n log-uniform in [2^14, 2^22], type uniform over "
                       "{Erdos-Renyi degree 16, R-MAT edgefactor 16 (0.57,0.19,0.19,0.05), RGG degree 32}, greedy MIS "
                       "each; %d batches per graph of 1,000-100,000 updates (log-uniform), endpoints in the "
                       "lowest-degree quarter; all requests submitted at once, interleaved across graphs"
                       % (len(graphs), args.steps),
           "name": "c5", "graphs": len(graphs), "requests": len(stats), "updates": int(total_up),
           "graph_entries": int(sum(len(g["nl"]) for g in graphs)),
           "served_updates_per_s": round(served_ups, 1), "plain_updates_per_s": round(total_up / plain_s, 1),
           "served_over_plain": round(served_ups / (total_up / plain_s), 3),
           "served_latency_p50_ms": round(pct(lat_all, 50), 3), "served_latency_p99_ms": round(pct(lat_all, 99), 3),
           "plain_p50_ms": round(pct(plain_lat, 50), 3), "plain_p99_ms": round(pct(plain_lat, 99), 3),
           "max_inflight_per_graph": args.inflight, "per_graph": per,
           "serving_kernel": "persistent %d: contended and small batches run their rounds (tiny batches whole) in "
                             "workgroups of one long-running kernel on a hardware queue of its own, the rest on "
                             "their graphs' streams" % args.persistent,
           "served_streams_only_updates_per_s": round(total_up / wall_streams, 1),
           "served_device_span_ms": round(dev_span, 3),
           "served_busy_fraction": round(busy_frac, 4),
           "served_mean_batches_in_progress": round(concurrency, 2),
           "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
           "open_loop": open_loop,
           "mis_plain_equals_served": True, "check_after_every_plain_batch": "independent, maximal, counts consistent",
           "device": dev_name}
    return {"metric": METRIC, "value": round(served_ups, 1), "unit": "edge updates/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 0, "ms_per_step": round(wall * 1e3 / len(stats), 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (seeded generators in misfit_amd/csrc/gen.cpp)", "config": cfg,
            "roofline": roof, "cpu_baseline": cpu}


def cpu_many(M, graphs, budget_s):
    """The oracle's literal engine on the first batch of each graph, smallest
    work first (batch size times graph entries), until the time budget is
    used; an item whose predicted time would overrun the budget is skipped.
    The GPU replays the same batches (plain) and must agree bit for bit."""
    from oracle import pyoracle as O
    threads = cpu_threads()
    items = sorted(range(len(graphs)), key=lambda gi: len(graphs[gi]["batches"][0]) * max(1, len(graphs[gi]["nl"])))
    spent, done, ups_done, lat, cost = 0.0, [], 0, [], None
    for gi in items:
        g = graphs[gi]
        b = g["batches"][0]
        work = len(b) * max(1, len(g["nl"]))
        if cost is not None and work * cost > 2.0 * (budget_s - spent):
            continue
        og = O.Graph.from_csr(g["ni"], g["nl"])
        m = g["mis0"].copy()
        t = time.perf_counter()
        og.process_batch(m, b["u"], b["v"], b["ins"], completion=True, literal=True, threads=threads)
        dt = time.perf_counter() - t
        del og
        spent += dt
        lat.append(dt)
        ups_done += len(b)
        cost = dt / work if cost is None else max(cost, dt / work)
        done.append((gi, m))
        if spent > budget_s:
            break
    same = True
    with M.Engine(0) as eng:
        for gi, m in done:
            g = graphs[gi]
            eng.load_csr(g["ni"], g["nl"])
            eng.set_mis(g["mis0"])
            eng.set_mode(M.MODE_PLAIN)
            eng.process_batch(g["batches"][0])
            same = same and bool(np.array_equal(eng.get_mis(), m))
    return {"value": ups_done / spent, "unit": "edge updates/s", "cores": threads, "kind": "port", **host_info(),
            "sample": "first batch of %d of the %d graphs, smallest first (%d updates), literal 2-hop rows "
                      "(OpenMP, %d threads) + sequential rule replay + completion; %.2f s"
                      % (len(done), len(graphs), ups_done, threads, spent),
            "p50_ms": round(pct([x * 1e3 for x in lat], 50), 3), "p99_ms": round(pct([x * 1e3 for x in lat], 99), 3),
            "bit_exact_vs_gpu": same}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=("c2", "c1", "c3", "c5", "compact"),
                    help="workload (default c2: R-MAT scale 22, the headline line)")
    ap.add_argument("--steps", type=int, default=None, help="timed batches (c5: batches per graph)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--graphs", type=int, default=64, help="c5: graphs held at once")
    ap.add_argument("--inflight", type=int, default=1, help="c5: batches on the device per graph")
    ap.add_argument("--workers", type=int, default=16, help="c5: server dispatcher threads")
    ap.add_argument("--persistent", type=int, default=1, choices=(0, 1, 2),
                    help="c5: serving kernel use (0 none, 1 contended and small batches, 2 every batch)")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--batch", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--served-only", action="store_true", help="skip the plain driver (profiling runs)")
    ap.add_argument("--no-stream-copy", action="store_true", help="skip the copy-bandwidth calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        # many graphs' streams share the hardware queues, each of which runs
        # its kernels in order: more queues (16, where fewer are set), less
        # waiting behind another graph's long batch (read once, when the HIP
        # runtime starts)
        try:
            hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        except ValueError:
            hwq = 4
        os.environ["GPU_MAX_HW_QUEUES"] = str(max(hwq, 16))
    copy = None if args.no_stream_copy else stream_copy_gbs()
    if copy:
        print("stream copy: %.1f GB/s" % copy["copy_gbs"], file=sys.stderr, flush=True)
    import misfit_amd as M

    steps_default = {"c2": 95, "c1": 10, "c3": 40, "c5": 4, "compact": 3}[args.config]
    warm_default = {"c2": 5, "c1": 2, "c3": 3, "c5": 0, "compact": 1}[args.config]
    args.steps = steps_default if args.steps is None else args.steps
    args.warmup = warm_default if args.warmup is None else args.warmup
    if args.config == "c1":
        W, K = args.warmup, args.steps
        ni, nl, mis0, u, v, gen_s = workload_c1(M, W, K)
        print("workload: n=%d m=%d generated in %.1fs" % (len(ni) - 1, len(nl), gen_s), file=sys.stderr, flush=True)
        wl = ("Erdos-Renyi n=100,000, 800,000 undirected edges, seed 42, greedy ascending-id MIS; 2,000-update "
              "batches of uniform vertex pairs (toggles, flags against the loaded graph)")
        print(json.dumps(single_graph_line(M, args, "c1", ni, nl, mis0, u, v, W, K, 2000, 8, copy, wl, "stream")))
        return
    if args.config == "c3":
        W, K = args.warmup, args.steps
        ni, nl, mis0, u, v, gen_s = workload_c3(M, W, K)
        print("workload: n=%d m=%d generated in %.1fs" % (len(ni) - 1, len(nl), gen_s), file=sys.stderr, flush=True)
        wl = ("2-D random geometric graph n=16,777,216, average degree 32, seed 11 (int64 offsets), greedy "
              "ascending-id MIS; 50,000-update batches with both endpoints drawn proportionally to (degree + 1)")
        print(json.dumps(single_graph_line(M, args, "c3", ni, nl, mis0, u, v, W, K, 50_000, 16, copy, wl, "stream")))
        return
    if args.config == "c5":
        print(json.dumps(many_graph_line(M, args, copy)))
        return
    if args.config == "compact":
        print(json.dumps(compact_line(M, args, copy)))
        return

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
        served_h = None
        if not args.served_only:
            served_h = run_served(M, eng, ni, nl, mis0, ups, W, K, B, staged=False)
            print("served, host hand-over: %.1f updates/s" % (K * B / served_h["wall_s"]), file=sys.stderr,
                  flush=True)
            if served_h["check"] != (0, 0, 0) or not np.array_equal(served["mis"], served_h["mis"]):
                raise SystemExit("served with host hand-over: final check failed or MIS differs")
        loads = None
        if plain is not None:
            loads = latency_at_load(M, eng, ni, nl, mis0, ups, B, W + K, K * B / plain["wall_s"])
        cpu = None
        if not args.no_cpu_baseline:
            # whole batches of the stream in order from the initial state, as
            # the reference's host path would run them
            cpu = cpu_stream(M, eng, ni, nl, mis0, ups, B, W + K, args.cpu_budget,
                             plain["bytes_all"] if plain is not None else None)
            print("cpu baseline: %.1f updates/s" % cpu["value"], file=sys.stderr, flush=True)
            if not cpu["bit_exact_vs_gpu"]:
                raise SystemExit("GPU differs from the oracle on the CPU-baseline sample")
    served_ups = K * B / served["wall_s"]
    roof = None
    if plain is not None and plain["bytes"]:
        tot_bytes = float(np.sum(plain["bytes"]))
        tot_s = served["region_ms"] / 1e3
        ach = tot_bytes / tot_s / 1e9
        traffic, traffic_src = committed_traffic(args)
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": KERNEL_NOTE,
                "algorithmic_bytes_per_batch": round(tot_bytes / K, 1),
                "algorithmic_bytes_per_update": round(tot_bytes / (K * B), 2),
                "kernel_ms_mean": round(served["region_ms"] / K, 4),
                "batch_device_clock_ms_mean": round(float(np.mean(served["dev_ms"])), 4),
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
        "served_phase_ms_mean": phase_means(served["stats"]),
        "served_submit_ms_mean": round(float(np.mean(served["submit_ms"])), 4),
        "served_host_handover_updates_per_s": round(K * B / served_h["wall_s"], 1) if served_h else None,
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
            "latency_note": "served_p50/p99: closed loop, 2 batches in flight, submit -> result including the "
                            "wait behind the batch ahead; open_loop: both drivers at the same offered load, "
                            "arrival -> result",
            "open_loop": loads,
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
