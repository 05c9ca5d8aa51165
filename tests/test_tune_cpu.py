"""The autotuner's search, its result file and its result check, on the CPU
with a stand-in for the timed run (the GPU measurement is covered by
tests/test_gpu_serve.py)."""
import json
import os

import numpy as np
import pytest

from misfit_amd import tune as T


def _fake(rates, mis_of=None, calls=None):
    """measure_fn: the rate is looked up by the knob settings (missing ones
    at 1.0); the MIS is fixed unless `mis_of` says otherwise."""
    good = np.array([1, 0, 1], np.uint8)

    def fn(env):
        key = tuple(sorted((k, v) for k, v in env.items() if v))
        if calls is not None:
            calls.append(key)
        mis = mis_of(key) if mis_of else good
        return rates.get(key, 1.0), mis
    return fn


SPACE = {"MISFIT_ROUND_MARGIN": ["", "1", "2"], "MISFIT_WINDOW_FACTOR": ["", "8", "32"]}


def test_descent_keeps_only_real_gains():
    rates = {
        (): 100.0,
        (("MISFIT_ROUND_MARGIN", "1"),): 100.5,        # under the 1% bar: not kept
        (("MISFIT_ROUND_MARGIN", "2"),): 110.0,        # kept
        (("MISFIT_ROUND_MARGIN", "2"), ("MISFIT_WINDOW_FACTOR", "8")): 90.0,
        (("MISFIT_ROUND_MARGIN", "2"), ("MISFIT_WINDOW_FACTOR", "32")): 120.0,  # kept
    }
    calls = []
    res = T.tune(_fake(rates, calls=calls), space=SPACE, repeats=1)
    assert res["env"] == {"MISFIT_ROUND_MARGIN": "2", "MISFIT_WINDOW_FACTOR": "32"}
    assert res["updates_per_s"] == 120.0 and res["default_updates_per_s"] == 100.0
    # the second knob is tried with the first at its best value
    assert (("MISFIT_ROUND_MARGIN", "2"), ("MISFIT_WINDOW_FACTOR", "8")) in calls
    assert len(res["trials"]) == 1 + 2 + 2


def test_defaults_win_when_nothing_beats_them():
    res = T.tune(_fake({(): 100.0}), space=SPACE, repeats=1)
    assert res["env"] == {} and res["updates_per_s"] == 100.0


def test_a_knob_that_changes_the_result_is_refused():
    bad = np.array([0, 1, 0], np.uint8)
    good = np.array([1, 0, 1], np.uint8)
    fn = _fake({}, mis_of=lambda key: bad if key == (("MISFIT_WINDOW_FACTOR", "8"),) else good)
    with pytest.raises(RuntimeError, match="changed the result"):
        T.tune(fn, space=SPACE, repeats=1)


def test_repeats_take_the_best_and_must_agree():
    seq = iter([(90.0, np.zeros(2, np.uint8)), (100.0, np.zeros(2, np.uint8))])
    res = T.tune(lambda env: next(seq), space={"MISFIT_ROUND_MARGIN": [""]}, repeats=2)
    assert res["updates_per_s"] == 100.0
    seq2 = iter([(90.0, np.zeros(2, np.uint8)), (100.0, np.ones(2, np.uint8))])
    with pytest.raises(RuntimeError, match="two different results"):
        T.tune(lambda env: next(seq2), space={"MISFIT_ROUND_MARGIN": [""]}, repeats=2)


def test_file_apply_and_exports(tmp_path, monkeypatch):
    for k in T.SPACE:
        # set, then delete: the knobs the tuner's apply() sets are undone after the test
        monkeypatch.setenv(k, "0")
        monkeypatch.delenv(k)
    monkeypatch.setenv("MISFIT_ROUND_WG_PER_CU", "7")  # not in the file: unset by apply
    res = {"env": {"MISFIT_ROUND_MARGIN": "2"}, "updates_per_s": 1.0, "default_updates_per_s": 1.0,
           "trials": []}
    p = str(tmp_path / "t.json")
    T.save(res, p)
    assert json.load(open(p))["env"] == {"MISFIT_ROUND_MARGIN": "2"}
    assert T.apply(p) == {"MISFIT_ROUND_MARGIN": "2"}
    assert os.environ["MISFIT_ROUND_MARGIN"] == "2"
    assert "MISFIT_ROUND_WG_PER_CU" not in os.environ
    lines = T.exports(p)
    assert "export MISFIT_ROUND_MARGIN=2" in lines and "unset MISFIT_WINDOW_FACTOR" in lines


def test_env_is_restored_after_a_trial(monkeypatch):
    monkeypatch.setenv("MISFIT_WINDOW_FACTOR", "16")
    monkeypatch.delenv("MISFIT_ROUND_MARGIN", raising=False)
    old = T._with_env({"MISFIT_ROUND_MARGIN": "3"})
    assert os.environ["MISFIT_ROUND_MARGIN"] == "3" and "MISFIT_WINDOW_FACTOR" not in os.environ
    T._restore(old)
    assert os.environ["MISFIT_WINDOW_FACTOR"] == "16" and "MISFIT_ROUND_MARGIN" not in os.environ
