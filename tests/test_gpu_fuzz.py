"""Randomised scenes, deterministic mode: for each of a fixed set of seeded
random boxes (face kinds, periodic pairs, volume gradients, pores, Inter
splits, grids of 0 to 3 dimensions, every tally kind, maxscat), the HIP
engine's per-bundle counts equal the FP64 checker's exactly and the tallies
agree to 1e-5, in a single solve and inside a batch; FP32 flight on the same
scenes runs every bundle and loses at most one in a million (one in a small solve)."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu
f = scenes.face
NCASES = 24


def random_case(k):
    rng = np.random.default_rng(1000 + k)
    dims = tuple(float(x) for x in rng.uniform(2e-8, 3e-7, 3))
    div = tuple(int(x) if rng.random() < 0.6 else 0 for x in rng.integers(1, 13, 3))
    kinds = [abi.SPEC, abi.DIFF, abi.PARTIAL]
    faces = [None] * 6
    pairs = []
    grad = [0.0, 0.0, 0.0]
    emit = False
    for ax, (lo, hi) in enumerate(((scenes.BAC, scenes.FRO), (scenes.LEF, scenes.RIG),
                                   (scenes.BOT, scenes.TOP))):
        r = rng.random()
        if r < 0.3:
            t = float(rng.uniform(0.5, 5.0))
            faces[lo], faces[hi] = f(abi.ISOT, t), f(abi.ISOT, -float(rng.uniform(0.0, 5.0)))
            emit = True
        elif r < 0.55:
            faces[lo], faces[hi] = f(abi.PERI), f(abi.PERI)
            pairs.append((lo, hi, ax))
            if rng.random() < 0.7:
                grad[ax] = float(rng.uniform(-2e6, 2e6))
                emit = emit or grad[ax] != 0.0
        else:
            for fi in (lo, hi):
                kk = kinds[int(rng.integers(0, 3))]
                faces[fi] = f(kk, 0.0, float(rng.uniform(0, 1)) if kk == abi.PARTIAL else 0.0)
    if not emit:
        faces[scenes.BAC] = f(abi.ISOT, 1.0)
        if faces[scenes.FRO][0] == abi.PERI:
            faces[scenes.FRO] = f(abi.ISOT, -1.0)
            pairs = [p for p in pairs if p[2] != 0]
            grad[0] = 0.0
    d = scenes.Domain(f"fuzz{k}")
    split = rng.random() < 0.25
    if split:
        # two boxes along x joined by an Inter pair
        half = (dims[0] / 2, dims[1], dims[2])
        fa = list(faces)
        fb = list(faces)
        fa[scenes.FRO] = f(abi.INTER)
        fb[scenes.BAC] = f(abi.INTER)
        dv = (max(div[0] // 2, 0), div[1], div[2])
        a = d.add_box((0, 0, 0), half, dv, fa, grad_t=tuple(grad))
        b = d.add_box((dims[0] / 2, 0, 0), half, dv, fb, grad_t=tuple(grad))
        d.pair(a, scenes.FRO, b, scenes.BAC)
        for lo, hi, ax in pairs:
            if ax == 0:
                d.pair(a, lo, b, hi, transl=(dims[0], 0.0, 0.0))
            else:
                t = [0.0, 0.0, 0.0]
                t[ax] = dims[ax]
                d.pair(a, lo, a, hi, transl=tuple(t))
                d.pair(b, lo, b, hi, transl=tuple(t))
    else:
        s = d.add_box((0, 0, 0), dims, div, faces, grad_t=tuple(grad))
        for lo, hi, ax in pairs:
            t = [0.0, 0.0, 0.0]
            t[ax] = dims[ax]
            d.pair(s, lo, s, hi, transl=tuple(t))
        if rng.random() < 0.3:
            r = 0.45 * min(dims[0], dims[1]) * float(rng.uniform(0.2, 0.9))
            d.add_pore(s, (dims[0] / 2, dims[1] / 2), r, kind=[abi.DIFF, abi.SPEC, abi.PARTIAL][int(rng.integers(0, 3))],
                       specularity=float(rng.uniform(0, 1)))
    tally = int(rng.integers(0, 5))
    maxscat = int(rng.integers(10, 120))
    size = int(rng.integers(1, 6)) if tally >= abi.CUMTEMP else 0
    nemit = int(rng.integers(4000, 15000))
    T = float(rng.uniform(150, 450))
    return materials.si_like(T), d, dict(nemit=nemit, tally=tally, maxscat=maxscat, size=size, seed=k + 1)


@pytest.fixture(scope="module")
def cases():
    return [random_case(k) for k in range(NCASES)]


@pytest.mark.parametrize("k", range(NCASES))
def test_fuzz_fp64_single(cases, k):
    mat, dom, p = cases[k]
    prob = scenes.problem(p["nemit"], tally=p["tally"], maxscat=p["maxscat"], size=p["size"], seed=p["seed"])
    print(f"case {k}: {dom.name} cells={len(dom.cells)} pores={len(dom.pores)} tally={p['tally']} "
          f"nemit={p['nemit']}", flush=True)
    with engine.Scene(mat, dom) as sc:
        fg, sg, cg = sc.solve(prob, counts=True)
    fo, so, co = oracle.solve(mat, dom, prob, counts=True)
    assert_counts_equal(cg, co)
    for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
        assert sg[key] == so[key], key
    assert np.all(rownorm_rel(fg, fo) < 1e-5)


def test_fuzz_batched_fp64_and_fp32(cases):
    """All random cases in one batch launch equal their own solves (FP64),
    and in FP32 flight every bundle of every case runs."""
    for prec in (abi.FP64, abi.FP32_FLIGHT):
        scs = [engine.Scene(m, d) for m, d, _ in cases]
        probs = [scenes.problem(p["nemit"], tally=p["tally"], maxscat=p["maxscat"], size=p["size"],
                                seed=p["seed"], precision=prec) for _, _, p in cases]
        fb, sb = engine.batch_solve(scs, probs)
        for k, (sc, pr) in enumerate(zip(scs, probs)):
            f1, s1, _ = sc.solve(pr)
            assert sb[k]["bundles"] == s1["bundles"] == sc.plan(pr).nemit
            if prec == abi.FP64:
                assert sb[k]["steps"] == s1["steps"], k
                assert np.all(rownorm_rel(fb[k], f1) < 1e-10), k
            else:   # FP32 rounding may lose a bundle on a face, rarely
                assert sb[k]["escapes"] <= max(1, sb[k]["bundles"] // 1_000_000), (k, sb[k]["escapes"])
                assert s1["escapes"] <= max(1, s1["bundles"] // 1_000_000), (k, s1["escapes"])
            assert np.all(np.isfinite(fb[k]))
        for sc in scs:
            sc.close()
