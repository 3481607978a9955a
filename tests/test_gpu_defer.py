"""Where a tally walk runs (mcg_tune.defer_walks: 1 in the wavefront's walk
phase, -1 inside the step, 0 by grid size) only schedules its additions: on
the same streams every bundle's counts are equal and the tallies agree up to
rounding (FP64: 1e-10, the shares' increments against the checker's
quotients; FP32 flight: 1e-5, FP32 against FP64 share products, in rows of
signed amounts that partly cancel), for 2-D and 3-D maps, single solves and
batches; the autotuner's choice keeps them too."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, octet, scenes
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu

STATS = ("bundles", "steps", "escapes", "absorbed", "truncated")
TOL = {abi.FP64: 1e-10, abi.FP32_FLIGHT: 1e-5}


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


def tuned(d, block=0):
    t = abi.Tune()
    t.defer_walks = d
    t.block = block
    return t


def octet_unit():
    return octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0)


SCENES = {
    "porous_32x32": lambda: scenes.porous_film(div=(32, 32, 0)),
    "porous_7x5x3": lambda: scenes.porous_film(div=(7, 5, 3)),
    "octet": octet_unit,
}


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_defer_walks_only_schedules(si300, name, precision):
    dom = SCENES[name]()
    prob = scenes.problem(40_000, maxscat=40, seed=8, precision=precision)
    print(f"defer_walks {name} precision {precision}", flush=True)
    out = {}
    with engine.Scene(si300, dom) as sc:
        for d in (0, 1, -1):
            sc.set_tune(tuned(d))
            out[d] = sc.solve(prob, counts=True)
    f0, s0, c0 = out[-1]
    for d in (0, 1):
        f, s, c = out[d]
        assert_counts_equal(c, c0)
        for k in STATS:
            assert s[k] == s0[k], (d, k)
        assert np.all(rownorm_rel(f, f0) < TOL[precision]), d


def test_defer_walks_batch(si300):
    """FP32 flight batch of grid scenes: deferred (1) and in-step (-1) walks."""
    doms = [scenes.porous_film(div=(16, 16, 0)), scenes.porous_film(div=(7, 5, 3)), octet_unit()]
    probs = [scenes.problem(20_000 + 5_000 * i, maxscat=40, seed=20 + i, precision=abi.FP32_FLIGHT)
             for i in range(len(doms))]
    sc = [engine.Scene(si300, d) for d in doms]
    try:
        fa, sa = engine.batch_solve(sc, probs, tuned(1))
        fb, sb = engine.batch_solve(sc, probs, tuned(-1))
        for i in range(len(probs)):
            for k in STATS:
                assert sa[i][k] == sb[i][k], (i, k)
            assert np.all(rownorm_rel(fa[i], fb[i]) < TOL[abi.FP32_FLIGHT]), i
    finally:
        for s in sc:
            s.close()


def test_autotune_grid_scene_keeps_results(si300):
    dom = scenes.porous_film(div=(32, 32, 0))
    prob = scenes.problem(100_000, maxscat=40, seed=9, precision=abi.FP32_FLIGHT)
    with engine.Scene(si300, dom) as sc:
        f0, s0, c0 = sc.solve(prob, counts=True)
        best, rate = sc.autotune(prob, sample_bundles=30_000)
        print(f"tuned: engine {best.engine} block {best.block} defer_walks {best.defer_walks}: {rate:.3g}",
              flush=True)
        assert rate > 0 and best.defer_walks in (-1, 0, 1)
        f1, s1, c1 = sc.solve(prob, counts=True)
    assert_counts_equal(c1, c0)
    for k in STATS:
        assert s1[k] == s0[k], k
    assert np.all(rownorm_rel(f1, f0) < 1e-5)
