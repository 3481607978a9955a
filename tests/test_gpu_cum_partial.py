"""The engine under the same identities that pin the checker's scatter-order
tallies and partially specular walls (tests/test_cum_partial.py), at sizes
where only such properties are checkable, and in both precisions.

* CumTemp / CumFlux prefix row k equals a Temp / Flux solve of the same
  streams truncated at maxscat = k * step + 1 (problem.cpp:433, 553-648):
  exact bundle for bundle in FP64, so the rows agree to summation order.
* A partial wall with p = 0 / p = 1 is DiffBoundary / SpecBoundary bundle for
  bundle (boundary.cpp:283-312), in FP64 and in FP32 flight.
* A partial wall draws specular with probability p (ballistic slab).
"""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes

pytestmark = pytest.mark.gpu

MAXSCAT, SIZE, STEP, MAXLOOP = 23, 5, 5, 10 ** 7


def _rows_close(a, b, rtol=1e-10):
    for r in range(b.shape[0]):
        scale = np.abs(b[r]).max()
        assert np.abs(a[r] - b[r]).max() <= rtol * scale + 1e-300, f"row {r}"


@pytest.mark.parametrize("kind,plain,width", [(abi.CUMTEMP, abi.TEMP, 1), (abi.CUMFLUX, abi.FLUX, 3)])
def test_cumulative_rows_are_truncated_solves(kind, plain, width):
    mat = materials.si_like(300.0)
    dom = scenes.bar(1e-6, 1e-7, 1e-7, 40, t_hot=1.0, t_cold=-1.0)
    n = 400_000
    print(f"case: cumulative kind={kind} n={n}", flush=True)
    with engine.Scene(mat, dom) as sc:
        cum, st, _ = sc.solve(scenes.problem(n, tally=kind, maxscat=MAXSCAT, maxloop=MAXLOOP, size=SIZE, seed=21))
        assert st["truncated"] > n // 20
        for k in range(SIZE + 1):
            ref, _, _ = sc.solve(scenes.problem(n, tally=plain, maxscat=min(k * STEP + 1, MAXSCAT),
                                                maxloop=MAXLOOP, seed=21))
            _rows_close(cum[width * k:width * (k + 1)], ref)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
@pytest.mark.parametrize("p,plain", [(0.0, abi.DIFF), (1.0, abi.SPEC)])
def test_partial_end_points_are_diff_and_spec(precision, p, plain):
    mat = materials.si_like(300.0)
    prob = scenes.problem(300_000, seed=31, precision=precision)
    print(f"case: partial end point p={p} precision={precision}", flush=True)
    with engine.Scene(mat, scenes.slab(1e-7, 10, side=abi.PARTIAL, specularity=p, lateral=2e-8)) as sc:
        fp, sp, cp = sc.solve(prob, counts=True)
    with engine.Scene(mat, scenes.slab(1e-7, 10, side=plain, lateral=2e-8)) as sc:
        fq, sq, cq = sc.solve(prob, counts=True)
    for name in ("nscat", "nbdry", "nstep", "fate"):
        assert np.array_equal(getattr(cp, name), getattr(cq, name)), name
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sp[key] == sq[key], key
    _rows_close(fp, fq, rtol=1e-9 if precision == abi.FP64 else 1e-6)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
def test_partial_specular_fraction(precision):
    p = 0.35
    mat = materials.grey(vel=1000.0, mfp=10.0)
    dom = scenes.slab(1e-7, 1, t_hot=1.0, t_cold=0.0, side=abi.PARTIAL, specularity=p, lateral=1e-8)
    print(f"case: partial fraction precision={precision}", flush=True)
    with engine.Scene(mat, dom) as sc:
        _, st, c = sc.solve(scenes.problem(200_000, seed=13, precision=precision), counts=True)
    assert st["escapes"] == 0 and st["absorbed"] == 200_000
    events = c.nbdry.astype(np.int64) - 1
    n = int(events.sum())
    frac = float((events - c.nscat.astype(np.int64)).sum()) / n
    assert abs(frac - p) < 4 * np.sqrt(p * (1 - p) / n), frac
