"""Writes tests/golden/oracle_vectors.npz: for each case of cases.py, the FP64
oracle's per-bundle counts (nscat, nbdry, nstep, fate), its solve statistics
and its normalised tallies.  Run from the repository root after a deliberate
change to the bundle streams or the physics:

    python -m tests.golden.make_vectors
"""
import os

import numpy as np

from oracle import oracle
from tests.golden.cases import cases

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_vectors.npz")


def main():
    arrays = {}
    for name, (mat, dom, prob) in cases().items():
        field, stats, counts = oracle.solve(mat, dom, prob, counts=True)
        arrays[f"{name}/field"] = field
        for c in ("nscat", "nbdry", "nstep", "fate"):
            arrays[f"{name}/{c}"] = getattr(counts, c)
        arrays[f"{name}/stats"] = np.array([stats[k] for k in ("bundles", "steps", "escapes", "absorbed",
                                                                 "truncated")], dtype=np.int64)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(arrays)} arrays")


if __name__ == "__main__":
    main()
