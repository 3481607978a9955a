"""Tallies too large for LDS stay in global memory, one private copy per XCD
(workgroup w adds into copy w % 8), summed into the first after the launch.
Deterministic mode equals the FP64 checker bundle by bundle with the tallies
within 1e-5 (a 128 x 128 Multi map of the porous film, 512 KiB; a 3-D map; a
1-D tally past the LDS cap of range adds); FP32 flight equals the batch
kernel's solve of the same problem, which tallies into one global copy, to
summation order, and a grid walk conserves the amounts."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


def bar_3d(div):
    """The bar with a 3-D map (scenes.bar's faces, divisions on every axis)."""
    sf = scenes.face(abi.DIFF)
    d = scenes.Domain("bar3d")
    d.add_box((0, 0, 0), (1e-6, 1e-7, 1e-7), div,
              [scenes.face(abi.ISOT, 10.0), sf, sf, scenes.face(abi.ISOT, -10.0), sf, sf])
    return d


CASES = {
    "porous_128x128": lambda: (scenes.porous_film(div=(128, 128, 0)), 20_000, 30),
    "bar_3d_40x16x16": lambda: (bar_3d((40, 16, 16)), 20_000, 60),
    "bar_1d_3000": lambda: (scenes.bar(1e-6, 1e-7, 1e-7, 3000, t_hot=10.0, t_cold=-10.0), 20_000, 60),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_global_tally_fp64_equals_checker(si300, name):
    dom, n, maxscat = CASES[name]()
    prob = scenes.problem(n, maxscat=maxscat, seed=3)
    print(f"global tally {name}", flush=True)
    with engine.Scene(si300, dom) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(si300, dom, prob, counts=True)
    assert_counts_equal(cg, co)
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[key] == so[key], key
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


@pytest.mark.parametrize("name", sorted(CASES))
def test_global_tally_fp32_equals_one_copy(si300, name):
    dom, n, maxscat = CASES[name]()
    prob = scenes.problem(20 * n, maxscat=maxscat, seed=4, precision=abi.FP32_FLIGHT)
    print(f"global tally fp32 {name}", flush=True)
    with engine.Scene(si300, dom) as sc:
        f1, s1, _ = sc.solve(prob)
        fb, sb = engine.batch_solve([sc], [prob])
    assert s1["steps"] == sb[0]["steps"] and s1["escapes"] == sb[0]["escapes"]
    assert np.all(rownorm_rel(f1, fb[0]) < 1e-9)
