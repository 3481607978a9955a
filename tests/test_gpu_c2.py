"""The batched workload at its full size: 256 independent slab solves
(thickness 10 nm-10 um, T_ref 100-500 K, lateral specularity U[0, 1],
1e4-1e6 bundles each, 5.4e7 in all) in one persistent launch and in one
replay of the captured graph.  Each solve stands for one solveField call
(main.cpp:146-179) of its own deck, so each must equal its own mcg_solve
(same streams; tallies differ by FP64 summation order only), and in FP64
the smallest eight equal the checker bundle for bundle."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, scenes, workloads
from oracle import oracle
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu


def test_slab_batch_256_equals_single_solves():
    cases = workloads.slab_batch(256, precision=abi.FP32_FLIGHT)
    assert sum(p.nemit for _, _, p in cases) > 5e7
    sc = [engine.Scene(m, d) for m, d, _ in cases]
    probs = [p for _, _, p in cases]
    try:
        print("c2: batch launch", flush=True)
        fb, sb = engine.batch_solve(sc, probs)
        print("c2: graph replay", flush=True)
        with engine.Batch(sc, probs) as g:
            fg, sg = g.run()
        print("c2: single solves", flush=True)
        for k, (s, p) in enumerate(zip(sc, probs)):
            f1, s1, _ = s.solve(p)
            for key in ("bundles", "steps", "escapes", "absorbed"):
                assert sb[k][key] == s1[key], (k, key)
                assert sg[k][key] == s1[key], (k, key)
            assert s1["escapes"] == 0, k
            assert np.all(rownorm_rel(fb[k], f1) < 1e-10), k
            assert np.all(rownorm_rel(fg[k], f1) < 1e-10), k
    finally:
        for s in sc:
            s.close()


def test_smallest_slabs_fp64_equal_checker():
    cases = workloads.slab_batch(256, precision=abi.FP64)
    small = sorted(range(256), key=lambda i: cases[i][2].nemit)[:8]
    sc = [engine.Scene(cases[i][0], cases[i][1]) for i in small]
    probs = [cases[i][2] for i in small]
    try:
        fb, sb = engine.batch_solve(sc, probs)
        for k, i in enumerate(small):
            m, d, p = cases[i]
            print(f"c2 fp64 slab {i}: {p.nemit} bundles", flush=True)
            fo, so, co = oracle.solve(m, d, p, counts=True)
            f1, s1, c1 = sc[k].solve(p, counts=True)
            assert_counts_equal(c1, co)
            for key in ("bundles", "steps", "escapes", "absorbed", "truncated"):
                assert sb[k][key] == so[key] and s1[key] == so[key], (i, key)
            assert np.all(rownorm_rel(fb[k], fo) < 1e-5), i
    finally:
        for s in sc:
            s.close()
