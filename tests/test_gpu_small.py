"""Tiny and ragged solves: bundle counts that leave most of a wavefront
empty or spill one lane into the next wavefront (nemit 1, 2, 63, 65, 1031;
FieldProblem's per-emitter rounding makes the bar's 2, 2, 62, 64, 1030 and
gives each of the octet's emitters at least one), on a
1-D bar, the porous film's 2-D map (FP64 walks deferred) and the octet
truss's 3-D grids.  FP64 equals the checker bundle for bundle with the
tallies within 1e-5; FP32 flight accounts for every bundle exactly once, in
a single solve and in a batch of the ragged sizes."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, octet, scenes
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 63, 65, 1031)
STATS = ("bundles", "steps", "escapes", "absorbed", "truncated")


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


SCENES = {
    "bar": lambda: scenes.bar(4e-7, 1e-7, 1e-7, 40, t_hot=10.0, t_cold=-10.0),
    "porous_64x64": lambda: scenes.porous_film(div=(64, 64, 0)),
    "octet": lambda: octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_small_fp64_equals_checker(si300, name):
    dom = SCENES[name]()
    with engine.Scene(si300, dom) as sc:
        for n in SIZES:
            prob = scenes.problem(n, maxscat=60, seed=40 + n)
            print(f"small fp64 {name} nemit={n}", flush=True)
            fg, sg, cg = sc.solve(prob, counts=True)
            fo, so, co = oracle.solve(si300, dom, prob, counts=True)
            assert_counts_equal(cg, co)
            for k in STATS:
                assert sg[k] == so[k], (n, k)
            assert np.all(rownorm_rel(fg, fo) < 1e-5), n


@pytest.mark.parametrize("name", sorted(SCENES))
def test_small_fp32_every_bundle_once(si300, name):
    dom = SCENES[name]()
    with engine.Scene(si300, dom) as sc:
        for n in SIZES:
            prob = scenes.problem(n, maxscat=60, seed=40 + n, precision=abi.FP32_FLIGHT)
            print(f"small fp32 {name} nemit={n}", flush=True)
            f, s, c = sc.solve(prob, counts=True)
            total = sc.plan(prob).nemit
            assert s["bundles"] == total
            assert s["escapes"] + s["absorbed"] + s["truncated"] == total
            fate, nstep = np.asarray(c.fate), np.asarray(c.nstep)
            assert s["steps"] == int(np.sum(nstep)) and np.all(nstep >= 1)
            assert np.sum(fate == abi.ABSORBED) == s["absorbed"]
            assert np.sum(fate == abi.ESCAPED) == s["escapes"]
            assert np.sum((fate == abi.MAXSCAT) | (fate == abi.MAXLOOP)) == s["truncated"]
            assert np.all(np.isfinite(f))


def test_small_batch(si300):
    """One batch launch of the ragged sizes over the three scenes (FP32 and
    FP64): each solve equals its own single solve."""
    doms = [SCENES[k]() for k in sorted(SCENES)]
    sc = [engine.Scene(si300, d) for d in doms]
    try:
        for prec in (abi.FP64, abi.FP32_FLIGHT):
            scs, probs = [], []
            for i, s in enumerate(sc):
                for n in SIZES:
                    scs.append(s)
                    probs.append(scenes.problem(n, maxscat=60, seed=70 + 7 * i + n, precision=prec))
            print(f"small batch precision {prec}: {len(probs)} solves", flush=True)
            fb, sb = engine.batch_solve(scs, probs)
            for k, (s, p) in enumerate(zip(scs, probs)):
                f1, s1, _ = s.solve(p)
                for key in STATS:
                    assert sb[k][key] == s1[key], (k, key)
                tol = 1e-10 if prec == abi.FP64 else 1e-5
                assert np.all(rownorm_rel(fb[k], f1) < tol), k
    finally:
        for s in sc:
            s.close()
