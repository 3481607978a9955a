"""The FP64 oracle reproduces the committed regression vectors of
tests/golden (per-bundle counts and statistics exactly, tallies to 1e-12):
a change to the bundle streams or to the oracle's arithmetic shows here on
the CPU before any GPU comparison."""
import os

import numpy as np
import pytest

from oracle import oracle
from tests.golden.cases import cases
from tests.helpers import rownorm_rel

VECTORS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")
STATS = ("bundles", "steps", "escapes", "absorbed", "truncated")


@pytest.fixture(scope="module")
def golden():
    with np.load(VECTORS) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(cases()))
def test_oracle_reproduces_vectors(golden, name):
    mat, dom, prob = cases()[name]
    field, stats, counts = oracle.solve(mat, dom, prob, counts=True)
    for c in ("nscat", "nbdry", "nstep", "fate"):
        np.testing.assert_array_equal(getattr(counts, c), golden[f"{name}/{c}"], err_msg=c)
    assert [stats[k] for k in STATS] == list(golden[f"{name}/stats"])
    assert np.all(rownorm_rel(field, golden[f"{name}/field"]) < 1e-12)
