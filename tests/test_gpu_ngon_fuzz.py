"""Randomised N-gon cells, deterministic mode: seeded random prisms (slanted
axes) and pyramids with 4 to 9-sided convex bases (jittered regular
polygons), random face kinds (isothermal, diffuse, specular, partially
specular), optional volume gradients, every tally kind and random maxscat.
The HIP engine's per-bundle counts equal the FP64 checker's and the tallies
agree to 1e-5; FP32 flight on the same cells runs every bundle and loses at
most one in a million (one in a small solve)."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel
from tests import ngon_scenes as ng

pytestmark = pytest.mark.gpu
f = scenes.face
NCASES = 16


def random_cell(k):
    rng = np.random.default_rng(5000 + k)
    n = int(rng.integers(4, 10))
    ang = 2 * np.pi * (np.arange(n) + rng.uniform(-0.3, 0.3, n)) / n - np.pi / 2
    r = rng.uniform(0.8, 1.2, n) * rng.uniform(0.5, 2.0) * ng.U
    base = np.stack([np.cos(ang) * r, np.sin(ang) * r], axis=1)
    base -= base[0]
    pyramid = rng.random() < 0.4
    top = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.5, 2.0)]) * ng.U
    if pyramid:
        top[:2] += base.mean(axis=0)
    nface = n + 1 if pyramid else n + 2
    kinds = [abi.SPEC, abi.DIFF, abi.PARTIAL]
    faces = []
    for i in range(nface):
        kk = kinds[int(rng.integers(0, 3))]
        faces.append(f(kk, 0.0, float(rng.uniform(0, 1)) if kk == abi.PARTIAL else 0.0))
    hot = int(rng.integers(0, nface))
    faces[hot] = f(abi.ISOT, float(rng.uniform(0.5, 3.0)))
    if rng.random() < 0.5:
        cold = (hot + 1 + int(rng.integers(0, nface - 1))) % nface
        faces[cold] = f(abi.ISOT, -float(rng.uniform(0.5, 3.0)))
    grad = tuple(rng.uniform(-2e6, 2e6, 3)) if rng.random() < 0.4 else (0.0, 0.0, 0.0)
    d = scenes.Domain(f"ngon{k}")
    d.add_cell(abi.PYRAMID if pyramid else abi.PRISM, (0.0, 0.0, 0.0), ng.ngon_mat(top, base), 0, faces,
               grad_t=grad)
    tally = int(rng.integers(0, 5))
    prob = scenes.problem(int(rng.integers(3000, 12000)), tally=tally, maxscat=int(rng.integers(5, 80)),
                          seed=int(rng.integers(1, 1 << 30)), size=int(rng.integers(1, 6)))
    return d, prob


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


@pytest.mark.parametrize("k", range(NCASES))
def test_random_ngon_fp64(si300, k):
    d, prob = random_cell(k)
    print(f"ngon fuzz {k}: {d.cells[0].cell} faces {d.cells[0].nface} tally {prob.tally}", flush=True)
    with engine.Scene(si300, d) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(si300, d, prob, counts=True)
    assert_counts_equal(cg, co)
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[key] == so[key], key
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


def test_random_ngon_fp32(si300):
    for k in range(NCASES):
        d, prob = random_cell(k)
        prob.precision = abi.FP32_FLIGHT
        with engine.Scene(si300, d) as sc:
            _, st, _ = sc.solve(prob)
        assert st["escapes"] <= max(1, st["bundles"] // 1_000_000), (k, st)
