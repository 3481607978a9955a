"""The HIP engine in deterministic mode reproduces the committed regression
vectors of tests/golden without the oracle library: per-bundle counts and
statistics exactly, tallies to 1e-5."""
import os

import numpy as np
import pytest

from montecarlocpp_amd import engine
from tests.golden.cases import cases
from tests.helpers import rownorm_rel

pytestmark = pytest.mark.gpu
VECTORS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")
STATS = ("bundles", "steps", "escapes", "absorbed", "truncated")


@pytest.fixture(scope="module")
def golden():
    with np.load(VECTORS) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(cases()))
def test_engine_reproduces_vectors(golden, name):
    mat, dom, prob = cases()[name]
    print(f"golden {name}", flush=True)
    with engine.Scene(mat, dom) as sc:
        field, stats, counts = sc.solve(prob, counts=True)
    for c in ("nscat", "nbdry", "nstep", "fate"):
        np.testing.assert_array_equal(getattr(counts, c), golden[f"{name}/{c}"], err_msg=c)
    assert [stats[k] for k in STATS] == list(golden[f"{name}/stats"])
    assert np.all(rownorm_rel(field, golden[f"{name}/field"]) < 1e-5)
