"""mcg_solve_static, the solve that allocates nothing (every per-solve
device buffer in the library's static device array), against mcg_solve
through the C ABI: two calls in a row, the second smaller, each equal to
mcg_solve on the same problem (same streams: counts exact, tallies to FP64
summation order), in both precisions and on a 2-D map; a problem too large
for the array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from tests.helpers import rownorm_rel

pytestmark = pytest.mark.gpu


def solve_static(sc, prob):
    info = sc.plan(prob)
    out = np.zeros(info.rows * info.cols)
    st = abi.Stats()
    engine.check(engine.lib().mcg_solve_static(sc.handle, C.byref(prob), abi.dptr(out), out.size, C.byref(st)))
    return out.reshape(info.cols, info.rows).T.copy(), st.as_dict()


CASES = [
    # (label, domain, bundles of the first call, of the second)
    ("bar", lambda: scenes.bar(1e-6, 1e-7, 1e-7, 100), 300_000, 40_000),
    ("porous", lambda: scenes.porous_film(div=(64, 64, 0)), 100_000, 10_000),
]


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
@pytest.mark.parametrize("label,make,n1,n2", CASES)
def test_static_equals_solve_twice(label, make, n1, n2, precision):
    mat = materials.si_like(300.0)
    tol = 1e-12 if precision == abi.FP64 else 1e-10
    with engine.Scene(mat, make()) as sc:
        for n, seed in ((n1, 5), (n2, 6)):
            prob = scenes.problem(n, maxscat=100, seed=seed, precision=precision)
            print(f"case: {label} n={n} precision={precision}", flush=True)
            fs, ss = solve_static(sc, prob)
            fg, sg, _ = sc.solve(prob)
            for k in ("bundles", "steps", "escapes", "absorbed", "truncated"):
                assert ss[k] == sg[k], k
            assert np.all(rownorm_rel(fs, fg) < tol), rownorm_rel(fs, fg)


def test_static_refuses_oversized_tally():
    mat = materials.si_like(300.0)
    dom = scenes.Domain("grid")
    sp = scenes.face(abi.SPEC)
    dom.add_box((0, 0, 0), (1e-7, 1e-7, 1e-7), (128, 128, 8),
                [scenes.face(abi.ISOT, 1.0), sp, sp, scenes.face(abi.ISOT, -1.0), sp, sp])
    with engine.Scene(mat, dom) as sc:
        with pytest.raises(engine.EngineError, match="bytes of device scratch"):
            solve_static(sc, scenes.problem(1000))
        # the ordinary solve still runs it
        f, st, _ = sc.solve(scenes.problem(1000))
        assert st["bundles"] == 1000
