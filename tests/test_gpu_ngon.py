"""Prisms and pyramids with N-gon bases on the GPU: a slanted pentagonal
prism with a volume source, a 4-gon pyramid, a regular 9-gon prism (11
faces: more than an FP32 step keeps in registers) and the 42-cell octet-truss
unit (OctetDomain) built from them.  Deterministic mode equals
the FP64 checker bundle by bundle; FP32 flight runs every bundle, loses at
most one in a million (one in a small solve) and agrees with FP64 within 4 standard errors over
seeds.  A rectangular 4-gon prism against the parallelepiped of the same box
gives the same heat flux within 4 standard errors, once the reference's prism
volume (half the box's: its volume loop leaves out the fan's last triangle)
is taken into account."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, octet, scenes
from oracle import oracle
from tests import ngon_scenes as ng
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu


def _nonagon():
    return ng.prism(ng.regular(9), np.array([0.1, -0.2, 1.2]) * ng.U, grad=(1e6, 0.0, 0.0))


SCENES = {
    "pentagon_prism": lambda: ng.prism(),
    "quad_pyramid": lambda: ng.pyramid(grad=(0.0, 1e6, 0.0)),
    "nonagon_prism": _nonagon,
    "octet": lambda: octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0),
}


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_ngon_fp64_equals_checker(si300, name):
    d = SCENES[name]()
    prob = scenes.problem(20000, maxscat=60, seed=7)
    print(f"ngon {name} fp64", flush=True)
    with engine.Scene(si300, d) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(si300, d, prob, counts=True)
    assert_counts_equal(cg, co)
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[key] == so[key], key
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_ngon_fp32_flight(si300, name):
    d = SCENES[name]()
    print(f"ngon {name} fp32", flush=True)
    q32, q64 = [], []
    with engine.Scene(si300, d) as sc:
        for seed in range(1, 9):
            for prec, out in ((abi.FP32_FLIGHT, q32), (abi.FP64, q64)):
                prob = scenes.problem(40000, maxscat=60, seed=seed, precision=prec)
                f, st, _ = sc.solve(prob)
                assert st["bundles"] == engine.plan(si300, d, prob)[0].nemit
                assert st["escapes"] <= max(1, st["bundles"] // 1_000_000), st["escapes"]
                out.append(f[1:4].sum())
    q32, q64 = np.array(q32), np.array(q64)
    se = np.sqrt(q32.var(ddof=1) / len(q32) + q64.var(ddof=1) / len(q64))
    assert abs(q32.mean() - q64.mean()) < 4 * se + 1e-12 * abs(q64.mean())


def test_rectangular_prism_is_the_box(si300):
    """Same box, same walls: the prism's one-cell tally is the box's times
    V_box / V_ref = 2 (FP32 flight: the prism is detected as a box too)."""
    p, b = ng.box_pair()
    print("ngon box pair", flush=True)
    qp, qb = [], []
    with engine.Scene(si300, p) as sp, engine.Scene(si300, b) as sb:
        for seed in range(1, 9):
            for sc, out in ((sp, qp), (sb, qb)):
                f, st, _ = sc.solve(scenes.problem(100000, maxscat=80, seed=seed,
                                                   precision=abi.FP32_FLIGHT))
                assert st["bundles"] == 100000 and st["escapes"] <= 1
                out.append(f[1, 0])
    qp, qb = np.array(qp) / 2.0, np.array(qb)
    se = np.sqrt(qp.var(ddof=1) / len(qp) + qb.var(ddof=1) / len(qb))
    assert abs(qp.mean() - qb.mean()) < 4 * se
    assert qb.mean() > 10 * se   # the flux is resolved, so the test has teeth
