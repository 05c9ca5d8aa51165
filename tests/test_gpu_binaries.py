"""The drop-in binaries on the bundled sample, end to end on the GPU.

`misfit` and `misfit_async` are run exactly as the reference's binaries are
(five positional arguments, src/misfit/misfit.cu:737-743): the final MIS file
they write under the working directory must hold the known answer the
oracle re-derives (tests/golden/sample_known_answer.json), and the input
.egr must be replaced by the compacted wide form the oracle's compaction
gives (core/flash_compactor.h:20-104), unless MISFIT_SKIP_COMPACTION=1.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT, SAMPLE
from tests import helpers as H

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "misfit_amd", "bin")


def _stage(tmp_path, narrow=False, engine_mod=None):
    d = tmp_path / "run"
    (d / "batches").mkdir(parents=True)
    for f in os.listdir(os.path.join(SAMPLE, "batches")):
        shutil.copy(os.path.join(SAMPLE, "batches", f), d / "batches" / f)
    shutil.copy(os.path.join(SAMPLE, "sample_mis.txt"), d / "sample_mis.txt")
    egr = d / "sample.egr"
    if narrow:
        ni, nl, _ = engine_mod.read_egr(os.path.join(SAMPLE, "sample.egr"))
        engine_mod.write_narrow_egr(str(egr), ni, nl)
    else:
        shutil.copy(os.path.join(SAMPLE, "sample.egr"), egr)
    return d


def _run(binary, d, extra_env=None):
    env = dict(os.environ)
    env.update(extra_env or {})
    args = [os.path.join(BIN, binary), "sample.egr", "sample_mis.txt", "batches", "5", "0"]
    print("running", binary, extra_env, flush=True)
    r = subprocess.run(args, cwd=str(d), env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _final_ids(d):
    p = d / "datasets" / "Final_MISs" / "sample_flash_final_mis.txt"
    return [int(x) for x in p.read_text().split()]


@pytest.mark.parametrize("binary", ["misfit", "misfit_async"])
@pytest.mark.parametrize("rule", ["completion", "literal"])
def test_binary_final_mis_and_compaction(oracle, engine_mod, tmp_path, binary, rule):
    golden = json.load(open(os.path.join(GOLDEN, "sample_known_answer.json")))[rule]
    d = _stage(tmp_path)
    env = {"MISFIT_RULE": "reference"} if rule == "literal" else {}
    _run(binary, d, env)
    assert _final_ids(d) == golden["final_ids"]
    # the input graph now holds the compacted graph, wide form
    g, mis0, batches = H.sample_inputs(oracle, SAMPLE)
    ni, nl = g.base_csr()
    _, _, og = H.oracle_replay(oracle, ni, nl, mis0, batches, completion=(rule == "completion"), literal=True)
    oni, onl = og.compact()
    eni, enl, wide = engine_mod.read_egr(str(d / "sample.egr"))
    assert wide
    assert len(enl) == golden["compacted_entries"]
    assert np.array_equal(eni, oni) and np.array_equal(enl, onl)


@pytest.mark.parametrize("binary", ["misfit", "misfit_async"])
def test_binary_narrow_input_and_skip_compaction(oracle, engine_mod, tmp_path, binary):
    golden = json.load(open(os.path.join(GOLDEN, "sample_known_answer.json")))["completion"]
    d = _stage(tmp_path, narrow=True, engine_mod=engine_mod)
    before = (d / "sample.egr").read_bytes()
    _run(binary, d, {"MISFIT_SKIP_COMPACTION": "1", "MISFIT_MODE": "plain"})
    assert _final_ids(d) == golden["final_ids"]
    assert (d / "sample.egr").read_bytes() == before


def test_run_script_sample(tmp_path):
    """run.sh --sample on the GPU: the bundled sample is replayed (on a copy of
    its graph) and the known final MIS is written under the working directory."""
    golden = json.load(open(os.path.join(GOLDEN, "sample_known_answer.json")))["completion"]
    sh = os.path.join(ROOT, "misfit_amd", "run.sh")
    before = open(os.path.join(SAMPLE, "sample.egr"), "rb").read()
    print("running run.sh --sample", flush=True)
    r = subprocess.run(["bash", sh, "--variant=misfit_async", "--sample"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ids = [int(x) for x in (tmp_path / "datasets" / "Final_MISs" / "sample_flash_final_mis.txt").read_text().split()]
    assert ids == golden["final_ids"]
    assert open(os.path.join(SAMPLE, "sample.egr"), "rb").read() == before
