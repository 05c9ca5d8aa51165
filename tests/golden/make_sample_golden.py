"""Writes tests/golden/sample_known_answer.json: the oracle's replay of the
reference's bundled sample (datasets/sample, copied into tests/golden/sample)
under the literal rule and with completion.  Run from the repository root:
    python tests/golden/make_sample_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as O  # noqa: E402

SAMPLE = os.path.join(ROOT, "tests", "golden", "sample")


def main():
    g = O.Graph.load_egr(os.path.join(SAMPLE, "sample.egr"))
    mis0, _ = O.read_mis(os.path.join(SAMPLE, "sample_mis.txt"), g.n)
    batches = [g.parse_batch_file(p) for p in O.batch_files_sorted(os.path.join(SAMPLE, "batches"))]
    ni, nl = g.base_csr()
    out = {"source": "oracle replay of datasets/sample (5 batches x 50 deletions)"}
    for completion in (False, True):
        gg = O.Graph.from_csr(ni, nl)
        m = mis0.copy()
        adm, adds, unc = [], [], []
        for (u, v, ins) in batches:
            st = gg.process_batch(m, u, v, ins, completion=completion, literal=True)
            adm.append(st["admit_del"] + st["admit_ins"])
            adds.append(st["completion_adds"])
            unc.append(gg.check(m)[1])
        out["completion" if completion else "literal"] = {
            "admissions": adm,
            "completion_adds": adds,
            "uncovered": unc,
            "final_card": int(m.sum()),
            "final_ids": [int(i) + 1 for i in np.nonzero(m)[0]],
            "compacted_entries": int(len(gg.compact()[1])),
        }
    path = os.path.join(ROOT, "tests", "golden", "sample_known_answer.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
