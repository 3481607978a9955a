"""OctetDomain (domain.cpp:591-1225): the 42-cell octet-truss unit, built
from 4-gon prisms and pyramids, triangular prisms and parallelepipeds.

Pinned independently of both host builders: every cell's convex hull is taken
from its vertices (scipy), the 52 check points of the reference each lie in
some cell, no cell's centre lies in another cell, and the struts carry the
only sources, 2 V |gradT| with V the strut's determinant.  Both builders
accept the 71 interface and 5 periodic pairs (antiparallel, coplanar, the
periodic map consistent) and agree on the plan; the checker's bundles never
leave the truss."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

from montecarlocpp_amd import abi, engine, materials, octet, scenes
from oracle import oracle

DIM = (6e-7, 1e-7, 1e-7, 1e-8)
DIV = (4, 2, 2, 1)


@pytest.fixture(scope="module")
def mat():
    return materials.si_like(300.0)


def vertices(kind, o, m):
    c = [m[:, k] for k in range(m.shape[1])]
    if kind == abi.PARALLELEPIPED:
        v = [o + i * c[0] + j * c[1] + k * c[2] for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    elif kind == abi.TRIPRISM:
        base = [o, o + c[0], o + c[1]]
        v = base + [p + c[2] for p in base]
    elif kind == abi.TETRAHEDRON:
        v = [o, o + c[0], o + c[1], o + c[2]]
    elif kind == abi.PRISM:
        base = [o] + [o + ck for ck in c[1:]]
        v = base + [p + c[0] for p in base]
    else:
        v = [o] + [o + ck for ck in c[1:]] + [o + c[0]]
    return np.array(v)


def hulls():
    table, pts = octet.cells(DIM, DIV)
    return [Delaunay(vertices(k, o, m)) for k, o, m, _, _ in table], table, pts


def test_checkpoints_inside_and_cells_disjoint():
    hs, table, pts = hulls()
    assert len(hs) == 42 and pts.shape == (52, 3)
    inside = np.array([[h.find_simplex(p) >= 0 for h in hs] for p in pts])
    assert inside.any(axis=1).all()
    for i, (k, o, m, _, _) in enumerate(table):
        cen = vertices(k, o, m).mean(axis=0)
        hit = [j for j, h in enumerate(hs) if h.find_simplex(cen) >= 0]
        assert hit == [i], (i, hit)


def test_plan_and_sources(mat):
    d = octet.octet(DIM, DIV, 1.0)
    assert len(d.cells) == 42 and len(d.pairs) == 76
    table, _ = octet.cells(DIM, DIV)
    s, t = DIM[0], DIM[3]
    r2 = np.sqrt(2.0)
    a = np.pi / 5 * (DIM[1] + DIM[2]) - (4 - np.pi) / 5 * t
    b = a / 4
    g = 1.0 / (s - 2 * a - 2 * (r2 + 1) * b - 2 * (2 + r2) * t)
    vols = [abs(np.linalg.det(m)) for k, o, m, dv, grad in table if grad]
    assert len(vols) == 10
    ncols = sum(int(np.prod(dv)) for k, o, m, dv, grad in table if grad)
    for lib in (oracle, engine):
        info, w, n = lib.plan(mat, d, scenes.problem(10000))
        assert info.nemitter == 10 and info.cols == ncols
        np.testing.assert_allclose(w, 2 * np.array(vols) * g * r2, rtol=1e-12)


def test_bundles_stay_in_the_truss(mat):
    d = octet.octet(DIM, DIV, 1.0)
    f, st = oracle.solve(mat, d, scenes.problem(4000, maxscat=100, seed=2))[:2]
    assert st["escapes"] == 0 and st["truncated"] == st["bundles"]
    assert np.isfinite(f).all()


def test_invalid_dimensions():
    with pytest.raises(ValueError, match="Invalid dimensions"):
        octet.octet((6e-7, 1e-7, 1e-7, 4e-8), DIV, 1.0)   # b <= t
    with pytest.raises(ValueError, match="Invalid dimensions"):
        octet.octet((3e-7, 1e-7, 1e-7, 1e-8), DIV, 1.0)   # period too short


def test_average_weights():
    """OctetDomain::average weighs the struts' first layer along their third
    axis by cell volume: with div (4, 2, 2, 1) that is 16 cells; a constant
    field averages to itself."""
    d = octet.octet(DIM, DIV, 1.0)
    vol = engine.cell_volumes(d)
    w = octet.average_weights(d, vol)
    assert np.count_nonzero(w) == 16 and np.all(w[w > 0] == vol[w > 0])
    field = np.vstack([np.full(vol.size, 3.0), np.arange(vol.size, dtype=float)])
    avg = octet.average(field, d, vol)
    assert avg.shape == (2, 1) and avg[0, 0] == pytest.approx(3.0, rel=1e-14)
