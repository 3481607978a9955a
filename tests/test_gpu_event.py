"""The event engine (MCG_ENGINE_EVENT: bundle state as structure of arrays in
HBM, one scatter-step per bundle per round, bundles regrouped by their next
event and cell kind between rounds) against the same references as the
history kernel: the committed regression vectors of every feature family in
deterministic FP64 (per-bundle counts and statistics exactly), the checker on
the workload slab, and the history engine in FP32 flight (same streams: the
counts exactly, tallies to rounding).  Every knob is exercised: bundles per
thread, rounds between regroupings, and a live set far below the bundle
count (many emission rounds, then the drain)."""
import os

import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes, workloads
from oracle import oracle
from tests.golden.cases import cases
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu
VECTORS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")
STATS = ("bundles", "steps", "escapes", "absorbed", "truncated")

SHAPES = {
    "default": dict(),
    "pt8_live512": dict(per_thread=8, live=512),
    "compact3_live300": dict(compact_every=3, live=300, per_thread=2),
}


def event_tune(**kw):
    t = abi.Tune()
    t.engine = abi.ENGINE_EVENT
    for k, v in kw.items():
        setattr(t, k, v)
    return t


@pytest.fixture(scope="module")
def golden():
    with np.load(VECTORS) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("name", sorted(cases()))
def test_event_engine_reproduces_vectors(golden, name, shape):
    mat, dom, prob = cases()[name]
    print(f"event golden {name} {shape}", flush=True)
    with engine.Scene(mat, dom) as sc:
        sc.set_tune(event_tune(**SHAPES[shape]))
        field, stats, counts = sc.solve(prob, counts=True)
    for c in ("nscat", "nbdry", "nstep", "fate"):
        np.testing.assert_array_equal(getattr(counts, c), golden[f"{name}/{c}"], err_msg=c)
    assert [stats[k] for k in STATS] == list(golden[f"{name}/stats"])
    assert np.all(rownorm_rel(field, golden[f"{name}/field"]) < 1e-5)


def test_event_engine_slab_equals_checker():
    mat, dom, prob = workloads.slab()
    with engine.Scene(mat, dom) as sc:
        sc.set_tune(event_tune(live=20_000))
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(mat, dom, prob, counts=True)
    assert_counts_equal(cg, co)
    for k in STATS:
        assert sg[k] == so[k], k
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


@pytest.mark.parametrize("label", ["bar", "porous"])
def test_event_engine_fp32_equals_history(label):
    if label == "bar":
        mat, dom = materials.si_like(300.0), scenes.bar(1e-6, 1e-7, 1e-7, 100)
        prob = scenes.problem(400_000, seed=3, precision=abi.FP32_FLIGHT)
    else:
        mat, dom, prob = workloads.porous_film(200_000, seed=4)
    print(f"event fp32 {label}", flush=True)
    with engine.Scene(mat, dom) as sc:
        fh, sh, ch = sc.solve(prob, counts=True)
        sc.set_tune(event_tune(live=65_536, per_thread=4))
        fe, se, ce = sc.solve(prob, counts=True)
    assert_counts_equal(ce, ch)
    for k in STATS:
        assert se[k] == sh[k], k
    assert np.all(rownorm_rel(fe, fh) < 1e-5)


def test_event_tune_rejects_bad_values():
    with engine.Scene(materials.si_like(300.0), scenes.slab(1e-7, 10)) as sc:
        for kw, msg in ((dict(per_thread=3), "per_thread"), (dict(engine=7), "unknown engine"),
                        (dict(compact_every=-1), "negative"), (dict(defer_walks=2), "defer_walks")):
            t = abi.Tune()
            for k, v in kw.items():
                setattr(t, k, v)
            with pytest.raises(engine.EngineError, match=msg):
                sc.set_tune(t)
