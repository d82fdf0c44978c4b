"""Writes quotient_c1.json: the oracle's multi-batch quotient passes
(oracle_rbsc_multibatch, a restatement of rbsc.cu:274-487 and
rbsc_aggressive.cu:487) on the 2^20-node R-MAT with 30% clones (seed 1001)
at batches of 50% and 25% of the lines, both stop rules: (nodes, done,
passes) and the SHA-256 of the canonical partition (uint32, little endian).
Run from the repository root: python tests/golden/make_quotient_c1.py"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import numpy as np  # noqa: E402

import oracle  # noqa: E402
from rbsc_amd import gen  # noqa: E402


def main():
    g = gen.rmat_2p20_clone30()
    src = g.sources()
    out = {"graph": "rmat_2p20_clone30 (seed 1001)", "n_nodes": g.n_nodes, "n_lines": g.n_edges,
           "cases": []}
    for frac in (2, 4):
        B = g.n_edges // frac
        for aggr in (False, True):
            nodes, done, passes, part = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, B,
                                                               aggressive=aggr, partition=True)
            h = hashlib.sha256(np.ascontiguousarray(part, "<u4").tobytes()).hexdigest()
            out["cases"].append({"max_batch_edges": B, "aggressive": aggr, "nodes": nodes,
                                 "done": done, "passes": passes, "partition_sha256": h})
            print(out["cases"][-1], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "quotient_c1.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
