"""The autotuner (mcg_autotune over both engines' knobs, mcg_autotune_batch
over the batch kernel's) only schedules: after it installs a shape, every
bundle's counts are the same and the tallies agree to FP64 summation order;
a batch run with the tuned shape equals the default batch."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes, workloads
from tests.helpers import assert_counts_equal, rownorm_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
def test_autotune_installs_shape_and_keeps_results(precision):
    mat, dom = materials.si_like(300.0), scenes.bar(4e-7, 1e-7, 1e-7, 40)
    prob = scenes.problem(300_000, seed=5, precision=precision)
    with engine.Scene(mat, dom) as sc:
        f0, s0, c0 = sc.solve(prob, counts=True)
        best, rate = sc.autotune(prob, sample_bundles=60_000)
        print(f"tuned: engine {best.engine} block {best.block} chunk {best.chunk} tables {best.lds_tables} "
              f"live {best.live} per_thread {best.per_thread} compact {best.compact_every}: {rate:.3g} steps/s",
              flush=True)
        assert rate > 0
        assert best.engine in (abi.ENGINE_HISTORY, abi.ENGINE_EVENT)
        f1, s1, c1 = sc.solve(prob, counts=True)
        # and the other engine, explicitly
        t = abi.Tune()
        t.engine = abi.ENGINE_EVENT if best.engine == abi.ENGINE_HISTORY else abi.ENGINE_HISTORY
        sc.set_tune(t)
        f2, s2, c2 = sc.solve(prob, counts=True)
    for c, s, f in ((c1, s1, f1), (c2, s2, f2)):
        assert_counts_equal(c, c0)
        for k in ("bundles", "steps", "escapes", "absorbed", "truncated"):
            assert s[k] == s0[k], k
        assert np.all(rownorm_rel(f, f0) < 1e-10)


def test_autotune_batch_shape_keeps_results():
    cases = workloads.slab_batch(24, precision=abi.FP32_FLIGHT)
    sc = [engine.Scene(m, d) for m, d, _ in cases]
    probs = [p for _, _, p in cases]
    try:
        best, rate = engine.autotune_batch(sc, probs, sample_fraction=0.2)
        print(f"batch tuned: block {best.block} chunk {best.chunk} tables {best.lds_tables}: {rate:.3g}", flush=True)
        assert rate > 0 and best.block in (128, 256, 512, 1024)
        f0, s0 = engine.batch_solve(sc, probs)
        f1, s1 = engine.batch_solve(sc, probs, best)
        for k in range(len(probs)):
            assert s1[k]["steps"] == s0[k]["steps"]
            assert np.all(rownorm_rel(f1[k], f0[k]) < 1e-10)
    finally:
        for s in sc:
            s.close()
