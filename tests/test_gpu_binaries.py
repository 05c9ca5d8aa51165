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


# The reference's report lines (src/misfit/misfit.cu:744-794,836-871,
# 931-960; src/misfit_async/misfit_async.cu:756-784,836-870,933-968;
# core/reporting.h:19-47), as patterns: the figures differ, the labels and
# their formatting do not.
_NUM = r"-?\d+\.\d{3}"
_SYNC_HEAD = [r"=+START=+", r"=+ MISFIT \(SYNC\) =+",
              r"Graph average degree: [\d.e+]+, MAX_CSR_ELEMENTS set to: \d+", r"Initial MIS cardinality: \d+"]
_ASYNC_HEAD = [r"=+START=+", r"=+ MISFIT_Async \(PIPELINED\) =+", r"=== MISFIT mis ===", r"Graph: \S+",
               r"  nodes=\d+ edges=\d+ avg_degree=\d+\.\d{2}", r"GPU: .+ \| discrete \(explicit-transfer path\)",
               r"MAX_CSR_ELEMENTS set to: \d+"]


def _batch_block(async_):
    head = [r"", r"=" * 60, r"Processing batch \d+/\d+ \| Edges: \d+"]
    if async_:
        head += [r"Batch \d+/\d+ \| edges=\d+ \| time=%sms \| clusters=\d+" % _NUM,
                 r"  Current MIS cardinality: \[Deferred for async execution\]",
                 r"=== BATCH TIMING BREAKDOWN \(algorithm-specific detail\) ==="]
    else:
        head += [r"  Current MIS cardinality: \[not computed per-batch\]", r"=== BATCH TIMING BREAKDOWN ==="]
    return head + [
        r"Total Batch Time: %s ms" % _NUM, r"", r"Component Breakdown:",
        r"  1\. CPU Conversion:        %s ms \(CPU\)" % _NUM,
        r"  2\. Host->Device Transfer: %s ms \(GPU\)" % _NUM,
        r"  3\. Cluster Processing:    %s ms \(GPU\)" % _NUM,
        r"       - Extract vertices: %s ms" % _NUM,
        r"       - Compute neighb\.:  %s ms" % _NUM,
        r"       - Vertex cluster\.:  %s ms" % _NUM,
        r"       - Assign edges:     %s ms" % _NUM,
        r"       - Clear assigns\.:   %s ms" % _NUM,
        r"  4\. GPU Process:           %s ms \(GPU\)" % _NUM,
        r"  5\. Device->Host Transfer: %s ms \(GPU\)" % _NUM,
        r"  6\. CPU Graph Update:      %s ms \(CPU\)" % _NUM,
        r"  7\. CPU Overhead:          %s ms \(CPU\)" % _NUM,
        r"  8\. Cluster count:         \d+",
        r"Finished MIS for batch \d+ -> slot [01]"]


_FINAL = [r"Saved final MIS to datasets/Final_MISs/sample_flash_final_mis\.txt", r"", r"=" * 80,
          r"=== FINAL RESULTS ===", r"=" * 80, r"", r"Graph Statistics:", r"  Initial MIS cardinality: \d+",
          r"  Final MIS cardinality:   \d+", r"  Cardinality change:      \d+", r"  Number of batches:       \d+",
          r"", r"Overall Timing Results:", r"  Total processing time:   %s milliseconds" % _NUM,
          r"  Average time per batch:  %s milliseconds" % _NUM, r"  Final D2H Transfer Time: %s milliseconds" % _NUM,
          r"  Processing rate:         \d+\.\d batches/second", r"", r"Average Component Times:",
          r"  Average Cluster Processing: %s ms" % _NUM, r"  Average TVB Time:           %s ms" % _NUM, r"",
          r"End-of-cycle compaction time \(NOT counted in per-batch numbers above\): %s ms" % _NUM,
          r"Peak program GPU memory: \d+\.\d{2} MB \(\d+\.\d% of \d+\.\d{2} MB total\)", r"=" * 80]
_SUMMARY = [r"", r"=== FINAL SUMMARY ===", r"Setup time:      %s ms" % _NUM, r"Processing time: %s ms" % _NUM,
            r"End-to-end time: %s ms" % _NUM, r"Initial MIS size: \d+", r"Final MIS size:   \d+"]


def _match_run(lines, i, pats):
    import re
    for k, p in enumerate(pats):
        assert i + k < len(lines) and re.fullmatch(p, lines[i + k]), \
            "line %d: %r does not match %r" % (i + k, lines[i + k] if i + k < len(lines) else None, p)
    return i + len(pats)


@pytest.mark.parametrize("binary", ["misfit", "misfit_async"])
def test_report_lines_match_reference(tmp_path, binary):
    """Scripts that read the reference's standard output read ours: the
    header, every batch block and the final block keep the reference's lines
    in the reference's order (this engine's own lines follow a batch's
    "Finished MIS" line and the final block)."""
    async_ = binary == "misfit_async"
    lines = _run(binary, _stage(tmp_path)).split("\n")
    i = _match_run(lines, 0, _ASYNC_HEAD if async_ else _SYNC_HEAD)
    block = _batch_block(async_)
    for b in range(5):
        i = _match_run(lines, i, block)
        assert ("Finished MIS for batch %d -> slot %d" % (b, b % 2)) == lines[i - 1]
        while i < len(lines) and lines[i].startswith("  "):   # this engine's detail lines
            i += 1
    i = _match_run(lines, i, _FINAL)
    if async_:
        i = _match_run(lines, i, _SUMMARY)
    assert lines[i] == "" and lines[i + 1] == "Engine summary:"


def test_engine_class_shim_on_the_sample(tmp_path):
    """The reference-shaped engine class (processBatchWithVertexClustering
    batch by batch, then launchBatch / syncAndCollect two in flight) ends the
    bundled sample with its known final MIS."""
    from tests.test_host_cpu import _build_shim_sample
    golden = json.load(open(os.path.join(GOLDEN, "sample_known_answer.json")))["completion"]
    exe = _build_shim_sample(tmp_path)
    print("running the engine class shim on the sample", flush=True)
    r = subprocess.run([exe, os.path.join(SAMPLE, "sample.egr"), os.path.join(SAMPLE, "sample_mis.txt"),
                        os.path.join(SAMPLE, "batches"), "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    runs = [[int(x) for x in line.split()] for line in r.stdout.strip().splitlines()]
    assert len(runs) == 2
    assert runs[0] == golden["final_ids"] and runs[1] == golden["final_ids"]
