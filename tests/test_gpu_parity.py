"""Deterministic-mode parity of the HIP engine against the FP64 CPU oracle.

Both run the same Philox stream per bundle, so in FP64 every bundle's
scatter count, boundary-hit count, step count and fate must be identical,
and the normalised tallies must agree within 1e-5 (row-norm relative; the
only differences are the order of FP64 additions).
"""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes, workloads
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu

TALLY_TOL = 1e-5


@pytest.fixture(scope="module")
def si300():
    return materials.si_like(300.0)


def run_both(mat, dom, prob):
    print(f"case: {dom.name} tally={prob.tally} prec={prob.precision} nemit={prob.nemit}", flush=True)
    with engine.Scene(mat, dom) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(mat, dom, prob, counts=True)
    return fg, sg, cg, fo, so, co


def check_det(fg, sg, cg, fo, so, co):
    assert_counts_equal(cg, co)
    for k in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[k] == so[k], k
    err = rownorm_rel(fg, fo)
    assert np.all(err < TALLY_TOL), err


def test_slab_workload_fp64(si300):
    """1-D cross-plane slab, 100 nm, isothermal walls +/-1 K, 1e5 bundles."""
    dom = scenes.slab(100e-9, 50)
    r = run_both(si300, dom, scenes.problem(100_000, seed=1))
    check_det(*r)


def test_bar_fp64(si300):
    """3-D bar 1 um x 100 nm x 100 nm with diffuse side walls."""
    dom = scenes.bar(1e-6, 1e-7, 1e-7, 100)
    r = run_both(si300, dom, scenes.problem(200_000, seed=7))
    check_det(*r)


@pytest.mark.parametrize("p", [0.0, 0.35, 1.0])
def test_partial_specularity_fp64(si300, p):
    dom = scenes.slab(300e-9, 60, side=abi.PARTIAL, specularity=p)
    r = run_both(si300, dom, scenes.problem(50_000, seed=11))
    check_det(*r)


@pytest.mark.parametrize("tally", [abi.TEMP, abi.FLUX, abi.MULTI])
def test_tally_kinds_fp64(si300, tally):
    dom = scenes.bar(4e-7, 1e-7, 5e-8, 40)
    r = run_both(si300, dom, scenes.problem(40_000, tally=tally, seed=5))
    check_det(*r)


@pytest.mark.parametrize("tally", [abi.CUMTEMP, abi.CUMFLUX])
def test_cumulative_kinds_fp64(si300, tally):
    dom = scenes.film((2e-7, 5e-8, 2e-7), (20, 0, 0), 1.0)
    r = run_both(si300, dom, scenes.problem(30_000, tally=tally, maxscat=41, size=8, seed=9))
    check_det(*r)


def test_bulk_periodic_volume_emitter(si300):
    """BulkDomain: periodic along x with the volumetric gradient emitter."""
    dom = scenes.bulk((2e-7, 1e-7, 1e-7), (16, 0, 0), 0.2)
    r = run_both(si300, dom, scenes.problem(30_000, maxscat=25, seed=3))
    check_det(*r)


def test_film_periodic(si300):
    dom = scenes.film((2e-7, 6e-8, 2e-7), (0, 12, 0), 0.2)
    r = run_both(si300, dom, scenes.problem(30_000, maxscat=30, seed=4))
    check_det(*r)


@pytest.mark.parametrize("div", [(12, 8, 0), (5, 4, 6)])
def test_grid_2d_3d_fp64(si300, div):
    """2-D / 3-D maps: the crossing walk must split segments as the ordered
    crossing map of the reference does."""
    d = scenes.Domain("box2d3d")
    f = scenes.face
    d.add_box((0, 0, 0), (3e-7, 2e-7, 1.5e-7), div,
              [f(abi.ISOT, 1.0), f(abi.DIFF), f(abi.SPEC), f(abi.ISOT, -1.0), f(abi.DIFF),
               f(abi.SPEC)])
    r = run_both(si300, d, scenes.problem(30_000, seed=21))
    check_det(*r)


def test_maxloop_truncation(si300):
    dom = scenes.film((2e-7, 6e-8, 2e-7), (10, 0, 0), 0.2)
    r = run_both(si300, dom, scenes.problem(20_000, maxscat=1000, maxloop=7, seed=8))
    check_det(*r)
    assert r[1]["truncated"] > 0


def test_bundle_subrange(si300):
    """A sub-range of bundles gives the same counts as the full run."""
    dom = scenes.slab(100e-9, 20)
    full = scenes.problem(20_000, seed=2)
    part = scenes.problem(20_000, seed=2, first=5000, count=7000)
    with engine.Scene(si300, dom) as sc:
        _, _, cf = sc.solve(full, counts=True)
        _, _, cp = sc.solve(part, counts=True)
    assert np.array_equal(cf.nscat[5000:12000], cp.nscat)
    assert np.array_equal(cf.nstep[5000:12000], cp.nstep)


def test_bundle_ids_past_32_bits(si300):
    """Bundle ids past 2^32 (a 2^33-bundle solve, run over the range that
    straddles 2^32, which is also where the second wall's bundles start): the
    ids' high words reach the streams and the emitter search, bundle for
    bundle as in the checker; FP32 flight runs the same range."""
    dom = scenes.slab(100e-9, 20)
    first = (1 << 32) - 3000
    prob = scenes.problem(1 << 33, seed=21, first=first, count=6000)
    r = run_both(si300, dom, prob)
    check_det(*r)
    assert r[1]["bundles"] == 6000
    p32 = scenes.problem(1 << 33, seed=21, first=first, count=6000, precision=abi.FP32_FLIGHT)
    with engine.Scene(si300, dom) as sc:
        _, st, _ = sc.solve(p32)
    assert st["bundles"] == 6000 and st["escapes"] <= 1


@pytest.mark.parametrize("name", ["jct", "tee", "tube"])
def test_assemblies_fp64(si300, name):
    """The reference's multi-box assemblies (domain.cpp:338-540): Inter faces
    with one and two partners, periodic pairs between different boxes."""
    dom = {"jct": lambda: scenes.jct((1e-7, 1e-7, 1e-7, 5e-8), (3, 3, 3, 2), 2e6 * 1e-7),
           "tee": lambda: scenes.tee((1e-7,) * 4 + (5e-8,), (3, 3, 3, 3, 0), 3e6 * 1e-7),
           "tube": lambda: scenes.tube((2e-7, 1e-7, 1e-7, 5e-8), (0, 4, 4, 4), 1e6 * 2e-7)}[name]()
    r = run_both(si300, dom, scenes.problem(30_000, maxscat=40, seed=17))
    check_det(*r)


@pytest.mark.parametrize("pore,walls", [(abi.DIFF, abi.DIFF), (abi.PARTIAL, abi.SPEC)])
def test_porous_film_fp64(si300, pore, walls):
    """Periodic porous film: a cylindrical pore in a 300 x 300 x 100 nm cell
    periodic in x (gradient source) and y, 64 x 64 map."""
    dom = scenes.porous_film(div=(64, 64, 0), pore=pore, walls=walls, specularity=0.4)
    r = run_both(si300, dom, scenes.problem(20_000, maxscat=40, seed=13))
    check_det(*r)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
@pytest.mark.parametrize("div", [(64, 64, 0), (7, 5, 3)])
def test_grid_walk_conserves_amounts(si300, precision, div):
    """A tally grid only splits each step's amounts over the cells it
    crosses: on the same streams, the volume-weighted sum of a grid tally
    equals the single-cell tally of the same scene (FP32 flight walks its
    2-D / 3-D tallies deferred, in the wavefront's walk phase, FP64 too in
    single solves; grids under 32 cells walk inside the step)."""
    grid = scenes.porous_film(div=div)
    whole = scenes.porous_film(div=(0, 0, 0))
    prob = scenes.problem(100_000, maxscat=60, seed=21, precision=precision)
    with engine.Scene(si300, grid) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    with engine.Scene(si300, whole) as sc:
        fw, sw, cw = sc.solve(prob, counts=True)
    assert_counts_equal(cg, cw)
    vg, vw = engine.cell_volumes(grid), engine.cell_volumes(whole)
    tot_g = fg @ vg
    tot_w = fw @ vw
    # deviational amounts carry both signs: bound the difference by the
    # summed magnitudes of the cells, not by the (partly cancelled) total
    scale = np.abs(fg) @ vg
    tol = 2e-6 if precision == abi.FP32_FLIGHT else 1e-12
    assert np.all(np.abs(tot_g - tot_w) <= tol * scale), (tot_g, tot_w, scale)


def test_porous_film_fp32_within_3_sigma(si300):
    """Production mode on the porous film: the effective conductivity from
    FP32 flight agrees with the FP64 checker on independent seeds."""
    dom = scenes.porous_film(div=(16, 16, 0))
    kg, ko = [], []
    with engine.Scene(si300, dom) as sc:
        for s in range(8):
            f = sc.solve(scenes.problem(20_000, maxscat=40, seed=s, precision=abi.FP32_FLIGHT))[0]
            kg.append(workloads.k_eff_porous(f, engine.cell_volumes(dom)))
    for s in range(8):
        f = oracle.solve(si300, dom, scenes.problem(20_000, maxscat=40, seed=100 + s))[0]
        ko.append(workloads.k_eff_porous(f, oracle.cell_volumes(dom)))
    se = np.sqrt(np.var(kg, ddof=1) / 8 + np.var(ko, ddof=1) / 8)
    assert abs(np.mean(kg) - np.mean(ko)) < 3 * se


def test_fp32_production_within_3_sigma(si300):
    """Production mode (FP32 flight, FP64 tallies) against the FP64 checker
    on independent seeds: heat flux within 3 standard errors, every T cell
    within 4."""
    dom = scenes.bar(4e-7, 1e-7, 1e-7, 20)
    g, o = [], []
    with engine.Scene(si300, dom) as sc:
        for s in range(8):
            g.append(sc.solve(scenes.problem(50_000, seed=s, precision=abi.FP32_FLIGHT))[0])
    for s in range(8):
        o.append(oracle.solve(si300, dom, scenes.problem(50_000, seed=100 + s))[0])
    g, o = np.array(g), np.array(o)
    qg, qo = g[:, 1].mean(axis=1), o[:, 1].mean(axis=1)
    se = np.sqrt(qg.var(ddof=1) / 8 + qo.var(ddof=1) / 8)
    assert abs(qg.mean() - qo.mean()) < 3 * se
    Tg, To = g[:, 0], o[:, 0]
    seT = np.sqrt(Tg.var(axis=0, ddof=1) / 8 + To.var(axis=0, ddof=1) / 8)
    assert np.all(np.abs(Tg.mean(axis=0) - To.mean(axis=0)) < 4 * seT)


def test_fp32_same_streams_close(si300):
    """FP32 flight follows the same random decisions as FP64.  On the
    specular-wall slab (no diffuse bounces to amplify rounding) almost every
    bundle keeps identical counts and the tallies agree to 1e-4; diffuse-wall
    geometries decorrelate the two precisions after tens of bounces, so there
    the comparison is statistical (test_fp32_production_within_3_sigma)."""
    dom = scenes.slab(1e-7, 50)
    with engine.Scene(si300, dom) as sc:
        f32, _, c32 = sc.solve(scenes.problem(200_000, seed=3, precision=abi.FP32_FLIGHT), counts=True)
        f64, _, c64 = sc.solve(scenes.problem(200_000, seed=3, precision=abi.FP64), counts=True)
    assert np.mean(c32.nscat == c64.nscat) > 0.99
    assert abs(f32[1].mean() / f64[1].mean() - 1) < 1e-4
    assert np.linalg.norm(f32[0] - f64[0]) / np.linalg.norm(f64[0]) < 1e-3


def test_fp32_large_solve_every_bundle_once(si300):
    """A solve larger than the end phase (two bundles per lane of the grid):
    next-bundle slots fill until the last ids, then claims shrink.  Every
    bundle id runs exactly once (each has a recorded fate and at least one
    step, and the steps add up to the solve's count), and the heat flux is
    that of FP64 on the same streams to 1e-3."""
    dom = scenes.bar(1e-6, 1e-7, 1e-7, 100, 10.0, -10.0)
    n = 1_500_000
    with engine.Scene(si300, dom) as sc:
        f32, s32, c32 = sc.solve(scenes.problem(n, seed=7, precision=abi.FP32_FLIGHT), counts=True)
        f64, s64, _ = sc.solve(scenes.problem(n, seed=7, precision=abi.FP64))
    assert s32["bundles"] == n
    assert np.all(c32.nstep >= 1)
    assert int(c32.nstep.astype(np.int64).sum()) == s32["steps"]
    assert np.all(c32.fate <= 3)
    assert s32["escapes"] <= max(1, n // 1_000_000)
    assert abs(f32[1].mean() / f64[1].mean() - 1) < 1e-3
