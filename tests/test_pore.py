"""The cylindrical-pore extension (a pore along z cut out of a box cell), on
the CPU: the rectangle-disk overlap behind the cut cells' volumes, the
emitter weight of the remaining material, construction errors, emission
outside the pore, and the checker's ray-cylinder flights and wall scattering
against independent restatements (numpy polynomial roots, reflection
geometry).  The engine's GPU path is compared with the checker bundle by
bundle in tests/test_gpu_parity.py."""
import math

import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from oracle import oracle

DIM = (300e-9, 300e-9, 100e-9)
R = 100e-9
C = (150e-9, 150e-9)


@pytest.fixture(scope="module")
def mat():
    return materials.si_like(300.0)


def brute_area(x0, x1, y0, y1, cx, cy, r, n=1500):
    """Midpoint-rule area of the rectangle inside the disk."""
    xs = x0 + (np.arange(n) + 0.5) * (x1 - x0) / n
    ys = y0 + (np.arange(n) + 0.5) * (y1 - y0) / n
    X, Y = np.meshgrid(xs, ys)
    return np.count_nonzero((X - cx) ** 2 + (Y - cy) ** 2 < r * r) * (x1 - x0) * (y1 - y0) / n ** 2


def test_disk_area_closed_forms():
    r = 1.3
    q = np.array([
        [-5, 5, -5, 5, 0.2, -0.1, r],          # whole disk: pi r^2
        [0.0, 5, -5, 5, 0.0, 0.0, r],          # right half
        [0.0, 5, 0.0, 5, 0.0, 0.0, r],         # quadrant
        [-0.3, 0.3, -0.2, 0.4, 0.0, 0.0, r],   # wholly inside
        [2.0, 3.0, 2.0, 3.0, 0.0, 0.0, r],     # wholly outside
        [0.5, 5, -5, 5, 0.0, 0.0, r],          # circular segment
    ], np.float64)
    a = oracle.disk_area(q)
    seg = r * r * math.acos(0.5 / r) - 0.5 * math.sqrt(r * r - 0.25)
    expect = [math.pi * r * r, math.pi * r * r / 2, math.pi * r * r / 4, 0.6 * 0.6, 0.0, seg]
    np.testing.assert_allclose(a, expect, rtol=1e-13, atol=1e-15)
    assert a[3] == 0.36 or a[3] == pytest.approx(0.36, rel=1e-15)
    assert a[4] == 0.0


def test_disk_area_brute_force():
    rng = np.random.default_rng(7)
    rows = []
    for _ in range(60):
        cx, cy = rng.uniform(-1, 1, 2)
        r = rng.uniform(0.2, 1.5)
        x0, y0 = rng.uniform(-2, 2, 2)
        rows.append([x0, x0 + rng.uniform(0.05, 2), y0, y0 + rng.uniform(0.05, 2), cx, cy, r])
    q = np.array(rows)
    a = oracle.disk_area(q)
    for row, got in zip(q, a):
        ref = brute_area(*row)
        cell = (row[1] - row[0]) * (row[3] - row[2])
        assert abs(got - ref) < 4e-3 * cell + 1e-12, (row, got, ref)


def test_engine_cell_volumes_match_checker():
    for d in (scenes.porous_film(DIM, R, (64, 64, 0)), scenes.porous_film(DIM, 0.9 * R, (5, 7, 3)),
              scenes.bar(1e-6, 1e-7, 1e-7, 100), scenes.film((2e-7, 6e-8, 2e-7), (0, 12, 3), 0.2)):
        np.testing.assert_allclose(engine.cell_volumes(d), oracle.cell_volumes(d), rtol=1e-14, atol=0)


def test_cell_volumes(mat):
    d = scenes.porous_film(DIM, R, (64, 64, 0))
    v = oracle.cell_volumes(d)
    assert v.size == 64 * 64
    box = DIM[0] * DIM[1] * DIM[2]
    assert v.sum() == pytest.approx(box - math.pi * R * R * DIM[2], rel=1e-12)
    grid = v.reshape(64, 64)   # [j, i]
    # symmetric about the pore's centre lines
    np.testing.assert_allclose(grid, grid[::-1, :], rtol=1e-11, atol=1e-40)
    np.testing.assert_allclose(grid, grid[:, ::-1], rtol=1e-11, atol=1e-40)
    # cells wholly inside have no material, cells wholly outside are whole
    h = DIM[0] / 64
    xc = (np.arange(64) + 0.5) * h
    far = np.hypot(np.abs(xc[None, :] - C[0]) + h / 2, np.abs(xc[:, None] - C[1]) + h / 2)
    near = np.hypot(np.maximum(np.abs(xc[None, :] - C[0]) - h / 2, 0),
                    np.maximum(np.abs(xc[:, None] - C[1]) - h / 2, 0))
    assert np.all(grid[far <= R] == 0.0)
    np.testing.assert_allclose(grid[near >= R], box / 4096, rtol=1e-14)
    cut = (far > R) & (near < R)
    assert cut.sum() > 100
    for j, i in zip(*np.nonzero(cut)[:2]):
        if (i + j) % 7:
            continue
        ref = brute_area(i * h, (i + 1) * h, j * h, (j + 1) * h, C[0], C[1], R, n=600)
        assert grid[j, i] == pytest.approx(box / 4096 * (1 - ref / h / h), abs=3e-3 * box / 4096)


def test_emitter_weight_and_plan(mat):
    d = scenes.porous_film(DIM, R, (64, 64, 0), dT=1.0)
    for lib in (oracle, engine):
        info, w, n = lib.plan(mat, d, scenes.problem(1000))
        vmat = DIM[0] * DIM[1] * DIM[2] - math.pi * R * R * DIM[2]
        np.testing.assert_allclose(w, [2 * vmat * (1.0 / DIM[0])], rtol=1e-14)
        assert list(n) == [1000] and info.cols == 4096 and info.rows == 4


def _err(lib, d, mat):
    with pytest.raises((oracle.OracleError, engine.EngineError)) as e:
        lib.plan(mat, d, scenes.problem(100))
    return str(e.value)


@pytest.mark.parametrize("lib", [oracle, engine], ids=["oracle", "engine"])
def test_pore_errors(lib, mat):
    f = scenes.face
    walls = [f(abi.SPEC)] * 6
    d = scenes.Domain()
    s = d.add_box((0, 0, 0), DIM, (4, 4, 0), walls)
    d.add_pore(s, C, 160e-9)
    assert "inside the cell" in _err(lib, d, mat)
    d = scenes.Domain()
    s = d.add_box((0, 0, 0), DIM, (4, 4, 0), walls)
    d.add_pore(s, C, -1e-9)
    assert "radius" in _err(lib, d, mat)
    d = scenes.Domain()
    s = d.add_box((0, 0, 0), DIM, (4, 4, 0), walls)
    d.add_pore(s, C, 5e-8)
    d.add_pore(s, C, 4e-8)
    assert "one pore" in _err(lib, d, mat)
    d = scenes.Domain()
    s = d.add_cell(abi.PARALLELEPIPED, (0, 0, 0),
                   np.array([[3e-7, 1e-7, 0], [0, 3e-7, 0], [0, 0, 1e-7]]), (4, 4, 0), walls)
    d.add_pore(s, C, 5e-8)
    assert "axis-aligned" in _err(lib, d, mat)
    d = scenes.Domain()
    s = d.add_box((0, 0, 0), DIM, (4, 4, 0), walls)
    d.add_pore(s, C, 5e-8, kind=abi.ISOT)
    assert "Diff, Spec or Partial" in _err(lib, d, mat)
    d = scenes.Domain()
    s = d.add_box((0, 0, 0), DIM, (4, 4, 0), walls)
    d.add_pore(s, C, 5e-8, kind=abi.PARTIAL, specularity=1.5)
    assert "specularity" in _err(lib, d, mat)
    # the hole through a face across the pore's axis must not open onto
    # another cell: Inter there, or Peri onto a different cell
    lo = list(walls)
    lo[scenes.TOP] = f(abi.INTER)
    hi = list(walls)
    hi[scenes.BOT] = f(abi.INTER)
    hi[scenes.BAC] = f(abi.ISOT, 1.0)
    d = scenes.Domain()
    a = d.add_box((0, 0, 0), DIM, (4, 4, 0), lo)
    b = d.add_box((0, 0, DIM[2]), DIM, (4, 4, 0), hi)
    d.pair(a, scenes.TOP, b, scenes.BOT)
    d.add_pore(a, C, 5e-8)
    assert "Inter" in _err(lib, d, mat)
    lo[scenes.TOP] = f(abi.PERI)
    hi[scenes.BOT] = f(abi.PERI)
    d = scenes.Domain()
    a = d.add_box((0, 0, 0), DIM, (4, 4, 0), lo)
    b = d.add_box((0, 0, DIM[2]), DIM, (4, 4, 0), hi)
    d.pair(a, scenes.TOP, b, scenes.BOT, transl=(0.0, 0.0, 0.0))
    d.add_pore(a, C, 5e-8)
    assert "within its cell" in _err(lib, d, mat)


def _cross_plane(T=1.0):
    """The pored box heated through its bottom face: isothermal walls across
    the pore's axis, so the face emitters carry the hole."""
    f = scenes.face
    walls = [f(abi.DIFF)] * 6
    walls[scenes.BOT] = f(abi.ISOT, T)
    walls[scenes.TOP] = f(abi.ISOT, 0.0)
    d = scenes.Domain("cross_plane")
    s = d.add_box((0, 0, 0), DIM, (0, 0, 8), walls)
    d.add_pore(s, C, R)
    return d


def test_face_emitter_weight_excludes_hole(mat):
    d = _cross_plane(2.0)
    for lib in (oracle, engine):
        info, w, n = lib.plan(mat, d, scenes.problem(1000))
        np.testing.assert_allclose(w, [(DIM[0] * DIM[1] - math.pi * R * R) * 2.0], rtol=1e-14)


def test_face_emission_outside_pore(mat):
    """Emission from a face across the pore is uniform over the face minus
    the disk, and every bundle of the scene stays in the material."""
    d = _cross_plane()
    n = 40000
    e = oracle.sample_emission(mat, d, scenes.problem(n, seed=7), n)
    assert np.all(e[:, 2] == 0.0)
    rr = np.hypot(e[:, 0] - C[0], e[:, 1] - C[1])
    assert rr.min() >= R
    area = DIM[0] * DIM[1] - math.pi * R * R
    p_ann = math.pi * ((1.4 * R) ** 2 - R * R) / area
    k = np.count_nonzero(rr < 1.4 * R)
    assert abs(k - n * p_ann) < 5 * math.sqrt(n * p_ann * (1 - p_ann))
    _, st = oracle.solve(mat, d, scenes.problem(5000, maxscat=50, seed=3))[:2]
    assert st["escapes"] == 0 and st["bundles"] == 5000


def test_emission_outside_pore_uniform(mat):
    d = scenes.porous_film(DIM, R, (64, 64, 0))
    n = 60000
    e = oracle.sample_emission(mat, d, scenes.problem(n, seed=11), n)
    rr = np.hypot(e[:, 0] - C[0], e[:, 1] - C[1])
    assert rr.min() >= R
    assert np.all((e[:, 0] >= 0) & (e[:, 0] <= DIM[0]) & (e[:, 1] >= 0) & (e[:, 1] <= DIM[1]))
    # uniform over the material: annulus R..1.4R versus the rest
    area = DIM[0] * DIM[1] - math.pi * R * R
    p_ann = math.pi * ((1.4 * R) ** 2 - R * R) / area
    k = np.count_nonzero(rr < 1.4 * R)
    assert abs(k - n * p_ann) < 5 * math.sqrt(n * p_ann * (1 - p_ann))
    # quadrant symmetry and uniform z
    q = np.count_nonzero((e[:, 0] < C[0]) & (e[:, 1] < C[1]))
    assert abs(q - n / 4) < 5 * math.sqrt(n * 3 / 16)
    assert abs(np.mean(e[:, 2]) / DIM[2] - 0.5) < 5 * math.sqrt(1 / 12 / n)


def _flights(mat, d, nb=40, seed=5):
    """(start, dir, dmin, hit, after) for every step after the first of nb
    traced bundles."""
    out = []
    prob = scenes.problem(2000, maxscat=60, seed=seed)
    for b in range(nb):
        t = oracle.trace(mat, d, prob, b)
        for k in range(1, len(t)):
            p = np.array([t[k - 1, 14], t[k - 1, 15], t[k - 1, 5]])
            out.append((p, t[k - 1, 6:9], t[k, 2], int(t[k, 1]), t[k]))
    return out


@pytest.mark.parametrize("kind", [abi.DIFF, abi.SPEC])
def test_flights_against_polynomial_roots(mat, kind):
    d = scenes.porous_film(DIM, R, (8, 8, 0), pore=kind)
    nhit = 0
    for p, u, dmin, hit, rec in _flights(mat, d):
        dx, dy = p[0] - C[0], p[1] - C[1]
        assert dx * dx + dy * dy >= (R * (1 - 1e-9)) ** 2
        # from outside the disk a ray can only enter while it heads towards
        # the axis; it enters at the smaller root of |p + t u|_xy = R
        roots = np.roots([u[0] ** 2 + u[1] ** 2, 2 * (dx * u[0] + dy * u[1]), dx * dx + dy * dy - R * R])
        real = [x.real for x in roots if abs(x.imag) < 1e-30]
        heading_in = dx * u[0] + dy * u[1] < 0
        tcyl = min(real) if real and heading_in else math.inf
        if hit == -3:
            nhit += 1
            assert dmin == pytest.approx(tcyl, rel=1e-9, abs=1e-18)
            a = rec[3:6]
            assert math.hypot(a[0] - C[0], a[1] - C[1]) == pytest.approx(R, rel=1e-12)
        else:
            assert dmin <= tcyl * (1 + 1e-9) + 1e-18
    assert nhit > 50


@pytest.mark.parametrize("kind", [abi.DIFF, abi.SPEC])
def test_pore_wall_scattering(mat, kind):
    d = scenes.porous_film(DIM, R, (8, 8, 0), pore=kind)
    cos_out = []
    for p, u, dmin, hit, rec in _flights(mat, d, nb=60):
        if hit != -3:
            continue
        a = rec[3:6]
        nrm = np.array([a[0] - C[0], a[1] - C[1], 0.0]) / R   # into the material
        v = rec[6:9]
        assert np.linalg.norm(v) == pytest.approx(1.0, abs=1e-14)
        if kind == abi.SPEC:
            refl = u - 2 * np.dot(u, nrm) * nrm
            np.testing.assert_allclose(v, refl, atol=1e-12)
        else:
            cos_out.append(np.dot(v, nrm))
    if kind == abi.DIFF:
        c = np.array(cos_out)
        assert c.min() > 0
        # Lambertian: cos theta ~ sqrt(U), mean 2/3
        assert np.mean(c) == pytest.approx(2 / 3, abs=5 / math.sqrt(18 * len(c)))


def test_vanishing_pore_reduces_to_plain_cell(mat):
    """A pore too small to be met leaves every trajectory of the periodic
    film bit-identical to the same cell without a pore (the pore code adds
    no draws and moves nothing), and the tallies equal to rounding."""
    with_pore = scenes.porous_film(DIM, 1e-15, (8, 8, 0))
    plain = scenes.porous_film(DIM, 1e-15, (8, 8, 0))
    plain.pores = []
    prob = scenes.problem(3000, maxscat=40, seed=9)
    fa, sa, ca = oracle.solve(mat, with_pore, prob, counts=True)
    fb, sb, cb = oracle.solve(mat, plain, prob, counts=True)
    for name in ("nscat", "nbdry", "nstep", "fate"):
        np.testing.assert_array_equal(getattr(ca, name), getattr(cb, name))
    np.testing.assert_allclose(fa, fb, rtol=1e-12, atol=1e-12 * np.abs(fb).max())
