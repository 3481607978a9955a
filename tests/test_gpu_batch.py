"""Batched solves (mcg_batch_solve): one persistent launch over many solves
with different materials, geometries, tally sizes and bundle counts must give
each solve exactly what its own mcg_solve gives (same streams; the tallies
differ only by FP64 summation order)."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes, workloads
from tests.helpers import rownorm_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
@pytest.mark.parametrize("tune", [None, abi.Tune(64, 1, 1000, 0, 0), abi.Tune(512, 1, 0, 1, -1)])
def test_batch_equals_individual_solves(precision, tune):
    cases = workloads.slab_batch(12, precision=precision)
    cases = [(m, d, scenes.problem(min(p.nemit, 60_000), seed=p.seed, precision=precision))
             for m, d, p in cases]
    cases.append((materials.si_like(300.0), scenes.bar(4e-7, 1e-7, 1e-7, 30),
                  scenes.problem(50_000, seed=99, precision=precision)))
    cases.append((materials.si_like(250.0), scenes.film((2e-7, 6e-8, 2e-7), (8, 0, 0), 0.3),
                  scenes.problem(20_000, tally=abi.CUMFLUX, size=5, maxscat=21, seed=98,
                                 precision=precision)))
    cases.append((materials.si_like(300.0), scenes.porous_film(div=(32, 32, 0)),
                  scenes.problem(20_000, maxscat=30, seed=97, precision=precision)))
    print("batch of", len(cases), "tune", tune and (tune.block, tune.chunk), flush=True)
    sc = [engine.Scene(m, d) for m, d, _ in cases]
    probs = [p for _, _, p in cases]
    fb, sb = engine.batch_solve(sc, probs, tune)
    for k, (s, p) in enumerate(zip(sc, probs)):
        f1, s1, _ = s.solve(p)
        assert sb[k]["steps"] == s1["steps"], k
        assert sb[k]["bundles"] == s1["bundles"]
        err = rownorm_rel(fb[k], f1)
        assert np.all(err < 1e-10), (k, err)
    for s in sc:
        s.close()


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
def test_graph_batch_replays_equal_eager(precision):
    """A batch captured into a HIP graph gives, on every replay, what the
    eager mcg_batch_solve gives, with eager work and synchronisations between
    the replays; new seeds on a replay equal an eager batch with those seeds."""
    cases = workloads.slab_batch(10, precision=precision)
    cases = [(m, d, scenes.problem(min(p.nemit, 40_000), seed=p.seed, precision=precision))
             for m, d, p in cases]
    cases.append((materials.si_like(300.0), scenes.porous_film(div=(16, 16, 0)),
                  scenes.problem(20_000, maxscat=30, seed=77, precision=precision)))
    sc = [engine.Scene(m, d) for m, d, _ in cases]
    probs = [p for _, _, p in cases]
    fe, se = engine.batch_solve(sc, probs)
    with engine.Batch(sc, probs) as b:
        for rep in range(3):
            fg, sg = b.run()
            print("replay", rep, flush=True)
            for k in range(len(probs)):
                assert sg[k]["steps"] == se[k]["steps"], (rep, k)
                assert np.all(rownorm_rel(fg[k], fe[k]) < 1e-10), (rep, k)
            sc[0].solve(probs[0])          # eager work between replays
        seeds = [1000 + k for k in range(len(probs))]
        fg, sg = b.run(seeds)
        for p, s in zip(probs, seeds):
            p.seed = s
        fe2, se2 = engine.batch_solve(sc, probs)
        for k in range(len(probs)):
            assert sg[k]["steps"] == se2[k]["steps"]
            assert np.all(rownorm_rel(fg[k], fe2[k]) < 1e-10)
    for s in sc:
        s.close()


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
def test_batch_of_one_scene_repeated(precision):
    """One scene listed several times in a batch (a stream of solves of one
    geometry, each with its own seed, and one seed twice): every entry gives
    what its own mcg_solve gives, and the repeated seed gives the same field."""
    with engine.Scene(materials.si_like(300.0), scenes.bar(4e-7, 1e-7, 1e-7, 30, t_hot=10.0, t_cold=-10.0)) as sc:
        seeds = [5, 6, 7, 5]
        probs = [scenes.problem(40_000, seed=s, precision=precision) for s in seeds]
        print("repeated scene", len(probs), "solves", flush=True)
        fb, sb = engine.batch_solve([sc] * len(probs), probs)
        for k, p in enumerate(probs):
            f1, s1, _ = sc.solve(p)
            assert sb[k]["steps"] == s1["steps"], k
            assert sb[k]["bundles"] == s1["bundles"] == 40_000
            assert np.all(rownorm_rel(fb[k], f1) < 1e-10), k
        assert sb[0]["steps"] == sb[3]["steps"]
        assert np.all(rownorm_rel(fb[0], fb[3]) < 1e-10)
        assert sb[0]["steps"] != sb[1]["steps"]
