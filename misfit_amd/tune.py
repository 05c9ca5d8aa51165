"""Autotuner for the engine's speed knobs on a caller's own graph and stream.

The engine reads five knobs when a context is created:
  MISFIT_ROUND_MARGIN       speculative rounds queued beyond the plan,
  MISFIT_WINDOW_FACTOR      edges a round speculates on (per expected commit),
  MISFIT_ROUND_WG_PER_CU    workgroups per CU that run a round,
  MISFIT_ROUND_THREADS      the round kernels' workgroup size (128/256/512),
  MISFIT_PU_DENSE_FRACTION  where the may-uncover scan switches from the
                            tagged vertices (sparse) to all vertices (dense).
They never change a result.  Their defaults come from measurements on the
R-MAT scale 22 stream.  A graph or
batch size far from that may prefer other values; this module finds them by
timing the queued driver on a short stream of the caller's own shape:

  coordinate descent from the defaults, one knob at a time, each candidate
  value timed with the others held at the best so far; a value is kept only
  when it beats the best by more than `min_gain`.  Every trial must end with
  the same MIS as the defaults (the knobs may not change results), or the
  tuner raises.

The result is a JSON file of environment settings: `apply(path)` sets them in
this process before contexts are created, and `exports(path)` prints them
for a shell that starts the `misfit` / `misfit_async` binaries.

  python -m misfit_amd.tune graph.egr --batch 10000 --batches 40 --out tune.json

The reference has no tuner: its launch shape is a constant
(THREADS_PER_BLOCK, src/misfit/misfit.cu:38, src/misfit_async/misfit_async.cu:33).
This one tunes the replacement engine only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

# candidate values per knob; "" leaves the knob unset (the engine's default)
SPACE = {
    "MISFIT_ROUND_MARGIN": ["", "1", "2", "3", "4"],
    "MISFIT_WINDOW_FACTOR": ["", "8", "12", "24", "32"],
    "MISFIT_ROUND_WG_PER_CU": ["", "2", "4"],
    "MISFIT_ROUND_THREADS": ["", "128", "256", "512"],
    "MISFIT_PU_DENSE_FRACTION": ["", "0.01", "0.1", "1"],
}


def _with_env(env):
    """Sets the knobs in `env` (unset for ""), returns the previous values."""
    old = {k: os.environ.get(k) for k in SPACE}
    for k in SPACE:
        v = env.get(k, "")
        if v:
            os.environ[k] = v
        else:
            os.environ.pop(k, None)
    return old


def _restore(old):
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def measure(M, ni, nl, mis0, ups, batch, warmup, steps, env, depth=2):
    """Queued-driver updates/s over `steps` batches after `warmup`, with the
    batches staged in device memory, in a context created under `env`.
    Returns (updates/s, final MIS)."""
    old = _with_env(env)
    try:
        with M.Engine(0) as eng:
            eng.load_csr(ni, nl)
            eng.set_mis(mis0)
            eng.set_mode(M.MODE_QUEUED)
            for b in range(warmup):
                eng.process_batch(ups[b * batch:(b + 1) * batch])
            hs = [eng.stage(ups[b * batch:(b + 1) * batch]) for b in range(warmup, warmup + steps)]
            eng.sync()
            t0 = time.perf_counter()
            tickets, nxt = [], 0
            for b in range(steps):
                while nxt < steps and nxt < b + depth:
                    tickets.append(eng.submit_staged(hs[nxt]))
                    nxt += 1
                eng.wait(tickets[b])
            eng.sync()
            dt = time.perf_counter() - t0
            mis = eng.get_mis()
            for h in hs:
                eng.release_staged(h)
            return steps * batch / dt, mis
    finally:
        _restore(old)


def tune(measure_fn, space=None, min_gain=0.01, repeats=2, log=None):
    """Coordinate descent over `space` (knob -> candidate values, "" = the
    default).  `measure_fn(env) -> (rate, mis)`; a trial's rate is the best of
    `repeats` runs.  Returns {"env", "updates_per_s", "default_updates_per_s",
    "trials"}."""
    space = SPACE if space is None else space
    trials = []

    def run(env):
        best, ref = 0.0, None
        for _ in range(repeats):
            r, mis = measure_fn(env)
            if ref is None:
                ref = mis
            elif not np.array_equal(ref, mis):
                raise RuntimeError("the same settings gave two different results: %s" % env)
            best = max(best, r)
        trials.append({"env": {k: v for k, v in env.items() if v}, "updates_per_s": round(best, 1)})
        if log:
            log("%s: %.4g updates/s" % (trials[-1]["env"] or "defaults", best))
        return best, ref

    cur = {k: "" for k in space}
    base, want = run(cur)
    best = base
    for k, values in space.items():
        for v in values:
            if v == cur[k]:
                continue
            env = dict(cur)
            env[k] = v
            r, mis = run(env)
            if not np.array_equal(mis, want):
                raise RuntimeError("knob setting %s changed the result" % {k: v})
            if r > best * (1.0 + min_gain):
                best, cur = r, env
    return {"env": {k: v for k, v in cur.items() if v}, "updates_per_s": round(best, 1),
            "default_updates_per_s": round(base, 1), "trials": trials}


def save(result, path):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)


def apply(path):
    """Sets the tuned knobs in this process's environment; contexts created
    afterwards use them.  Knobs the file leaves at the default are unset."""
    env = load(path)["env"]
    for k in SPACE:
        if k in env:
            os.environ[k] = str(env[k])
        else:
            os.environ.pop(k, None)
    return env


def exports(path):
    """Shell lines that set the tuned knobs."""
    env = load(path)["env"]
    return ["export %s=%s" % (k, env[k]) if k in env else "unset %s" % k for k in SPACE]


def main(argv=None):
    import misfit_amd as M
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("graph", help=".egr file")
    ap.add_argument("--batch", type=int, default=10000, help="updates per batch")
    ap.add_argument("--batches", type=int, default=40, help="timed batches per trial")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--kind", choices=["mixed", "uniform", "preferential", "low-degree"], default="mixed",
                    help="update stream drawn on the graph (as misfit_egr batches)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="misfit_tune.json")
    a = ap.parse_args(argv)
    ni, nl, _ = M.read_egr(a.graph)
    kind = {"mixed": M.MIXED_EXISTING, "uniform": M.UNIFORM_PAIRS, "preferential": M.PREFERENTIAL,
            "low-degree": M.LOW_DEGREE}[a.kind]
    nb = a.warmup + a.batches
    u, v = M.gen_updates(kind, ni, nl, nb * a.batch, a.seed)
    mis0 = M.greedy_mis(ni, nl)
    with M.Engine(0) as eng:
        eng.load_csr(ni, nl)
        ups = M.make_updates(u, v, eng.flags(u, v))
    res = tune(lambda env: measure(M, ni, nl, mis0, ups, a.batch, a.warmup, a.batches, env),
               log=lambda s: print(s, file=sys.stderr, flush=True))
    save(res, a.out)
    print(json.dumps({"env": res["env"], "updates_per_s": res["updates_per_s"],
                      "default_updates_per_s": res["default_updates_per_s"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
