"""Cells other than axis-aligned boxes on the GPU: a triangular prism, a
tetrahedron with a volume source and a skewed parallelepiped, each with an
isothermal face and diffuse / specular / partial walls and a 1-D to 3-D grid.
Deterministic mode equals the FP64 checker bundle by bundle (counts exact,
tallies to 1e-5); FP32 flight, which advects such cells face by face rather
than axis by axis, runs every bundle, loses at most one in a million (one in a small solve), and its
heat flux agrees with FP64 within 4 standard errors over seeds."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu

CELLS = {
    "prism": (abi.TRIPRISM, [[2e-7, 0, 0], [0, 1.5e-7, 0], [0, 0, 1e-7]], (5, 4, 3), (0.0, 0.0, 0.0)),
    "tetra": (abi.TETRAHEDRON, [[2e-7, 0, 0], [0.5e-7, 1.5e-7, 0], [0.3e-7, 0.2e-7, 1e-7]], (4, 4, 4),
              (2e6, 0.0, 0.0)),
    "skew": (abi.PARALLELEPIPED, [[2e-7, 0, 0], [0.6e-7, 1.5e-7, 0], [0, 0.3e-7, 1e-7]], (6, 0, 0),
             (0.0, 0.0, 0.0)),
}


def cell_domain(name):
    kind, m, div, grad = CELLS[name]
    nf = abi.NFACE[kind]
    f = scenes.face
    faces = [f(abi.ISOT, 1.0)] + [f((abi.DIFF, abi.SPEC, abi.PARTIAL)[i % 3], 0.0, 0.4 if i % 3 == 2 else 0.0)
                                  for i in range(1, nf)]
    d = scenes.Domain(name)
    d.add_cell(kind, (0.0, 0.0, 0.0), np.array(m), div, faces, grad_t=grad)
    return d


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


@pytest.mark.parametrize("name", sorted(CELLS))
def test_cell_fp64_equals_checker(si300, name):
    d = cell_domain(name)
    prob = scenes.problem(20000, maxscat=60, seed=5)
    print(f"cell {name} fp64", flush=True)
    with engine.Scene(si300, d) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(si300, d, prob, counts=True)
    assert_counts_equal(cg, co)
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[key] == so[key], key
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


@pytest.mark.parametrize("name", sorted(CELLS))
def test_cell_fp32_flight(si300, name):
    d = cell_domain(name)
    print(f"cell {name} fp32", flush=True)
    q32, q64 = [], []
    with engine.Scene(si300, d) as sc:
        for seed in range(1, 9):
            for prec, out in ((abi.FP32_FLIGHT, q32), (abi.FP64, q64)):
                f, st, _ = sc.solve(scenes.problem(40000, maxscat=60, seed=seed, precision=prec))
                assert st["bundles"] == 40000
                assert st["escapes"] <= max(1, st["bundles"] // 1_000_000), st["escapes"]
                out.append(f[1].sum())
    q32, q64 = np.array(q32), np.array(q64)
    se = np.sqrt(q32.var(ddof=1) / len(q32) + q64.var(ddof=1) / len(q64))
    assert abs(q32.mean() - q64.mean()) < 4 * se + 1e-12 * abs(q64.mean())
