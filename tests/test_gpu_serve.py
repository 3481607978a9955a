"""Continuous-batching server: every request's result is the same alone, in
a mix with other requests, and as a plain mcg_solve; admission respects the
device-memory budget; poll / cancel semantics."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, engine, materials, scenes, serve
from tests.helpers import rownorm_rel

pytestmark = pytest.mark.gpu


def drain(sv, ids):
    res = {}
    for _ in range(100000):
        sv.step()
        for rid in sv.completed():
            res[rid] = sv.poll(rid)
        if len(res) == len(ids):
            return res
    raise AssertionError("server did not finish")


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_FLIGHT])
def test_alone_equals_mixed_equals_solve(precision):
    fac = serve.RequestFactory(nmin=2e4, nmax=3e5, precision=precision)
    reqs = [fac.make(r) for r in range(1, 13)]
    target = reqs[3]
    print("target family", target[0], "n", target[3].nemit, flush=True)
    with engine.Server(step_bundles=50_000) as sv:
        a = sv.submit(*target[1:])
        alone = drain(sv, [a])[a]
    with engine.Server(step_bundles=70_000) as sv:
        ids = [sv.submit(*r[1:]) for r in reqs]
        mixed = drain(sv, ids)
    m = mixed[ids[3]]
    with engine.Scene(target[1], target[2]) as sc:
        f1, s1, _ = sc.solve(target[3])
    for got, st in ((alone[0], alone[1]), (m[0], m[1])):
        assert st["steps"] == s1["steps"] and st["bundles"] == s1["bundles"]
        assert np.all(rownorm_rel(got, f1) < 1e-10)
    # every other request of the mix also equals its own solve
    for rid, r in zip(ids, reqs):
        with engine.Scene(r[1], r[2]) as sc:
            f, st1, _ = sc.solve(r[3])
        assert mixed[rid][1]["steps"] == st1["steps"]
        assert np.all(rownorm_rel(mixed[rid][0], f) < 1e-10)


def test_budget_admission_and_cancel():
    fac = serve.RequestFactory(nmin=1e4, nmax=1e5, families=("slab", "bar", "random_slab"))
    reqs = [fac.make(r) for r in range(1, 7)]
    # each of these requests holds ~125 KB of device memory: a 200 KB budget
    # admits one
    with engine.Server(budget_bytes=200_000, step_bundles=20_000) as sv:
        ids = [sv.submit(*r[1:]) for r in reqs]
        a, q, _ = sv.step()
        assert a == 1 and q == 5
        info = sv.info()
        assert info["bytes_used"] <= info["bytes_budget"]
        sv.cancel(ids[-1])
        res = drain(sv, ids[:-1])
        assert len(res) == 5
        with pytest.raises(engine.EngineError) as e:
            sv.poll(ids[-1])
        assert e.value.code == abi.MCG_ENOTFOUND
    with engine.Server(budget_bytes=1000) as sv:
        with pytest.raises(engine.EngineError, match="budget"):
            sv.submit(*reqs[0][1:])


def test_open_loop_latency_report():
    out = serve.run(lam=200.0, nreq=40, factory=serve.RequestFactory(nmin=1e4, nmax=1e6))
    assert out["requests"] == 40
    assert 0 < out["latency_p50_s"] <= out["latency_p99_s"]
    assert out["scatter_steps_per_s"] > 0


def test_served_fp64_requests_equal_checker():
    """Requests served in FP64 (deterministic mode), mixed in one server,
    equal the CPU checker's solves of the same problems."""
    from oracle import oracle
    fac = serve.RequestFactory(nmin=1e4, nmax=4e4, precision=abi.FP64)
    reqs = [fac.make(r) for r in range(1, 9)]
    print("families", [r[0] for r in reqs], flush=True)
    with engine.Server(step_bundles=60_000) as sv:
        ids = [sv.submit(*r[1:]) for r in reqs]
        res = drain(sv, ids)
    for rid, (fam, mat, dom, prob) in zip(ids, reqs):
        fo, so, _ = oracle.solve(mat, dom, prob)
        assert res[rid][1]["steps"] == so["steps"], fam
        assert np.all(rownorm_rel(res[rid][0], fo) < 1e-5), fam


def test_rerun_sampling_reports_errors():
    fac = serve.RequestFactory(nmin=1e4, nmax=1e5, families=("slab", "bar", "random_slab"))
    out = serve.run(lam=400.0, nreq=20, factory=fac, rerun_fraction=0.1)
    errs = out["fp64_rerun_errors"]
    assert len(errs) == 2
    for e in errs:
        # FP32 flight and its FP64 rerun share the streams: their difference
        # is at most the solve's own noise, a few standard errors
        assert 0 <= e["heat_flux"] < 0.2 and 0 <= e["temperature"] < 0.2
        assert e["heat_flux_in_se"] < 4 and e["temperature_in_se"] < 4, e


def test_failed_request_reported():
    """A request is built on the server's host threads after mcg_server_submit
    returns: one whose problem cannot be solved (a closed box with no source)
    is completed with its error, which poll returns, and the requests around
    it are served as usual."""
    fac = serve.RequestFactory(nmin=1e4, nmax=3e4, families=("slab", "bar"))
    good = [fac.make(r) for r in range(1, 4)]
    sp = scenes.face(abi.SPEC)
    closed = scenes.Domain("closed")
    closed.add_box((0, 0, 0), (1e-7, 1e-7, 1e-7), (0, 0, 0), [sp] * 6)
    with engine.Server(step_bundles=40_000) as sv:
        ids = [sv.submit(*good[0][1:]), sv.submit(good[1][1], closed, scenes.problem(1000, seed=2)),
               sv.submit(*good[2][1:])]
        done, failed = {}, []
        for _ in range(100000):
            sv.step()
            for rid in sv.completed():
                if rid == ids[1]:
                    with pytest.raises(engine.EngineError, match="No phonons to emit"):
                        sv.poll(rid)
                    failed.append(rid)
                else:
                    done[rid] = sv.poll(rid)
            if len(done) + len(failed) == 3:
                break
        assert failed == [ids[1]] and sorted(done) == sorted([ids[0], ids[2]])
        with pytest.raises(engine.EngineError) as e:
            sv.poll(ids[1])
        assert e.value.code == abi.MCG_ENOTFOUND
    for rid, r in ((ids[0], good[0]), (ids[2], good[2])):
        with engine.Scene(r[1], r[2]) as sc:
            f, st, _ = sc.solve(r[3])
        assert done[rid][1]["steps"] == st["steps"]
        assert np.all(rownorm_rel(done[rid][0], f) < 1e-10)


def test_cancel_while_building():
    """Cancelling requests right after submission (their material tables may
    still be building on the server's threads) leaves the others to finish
    with their results, and the cancelled ids unknown."""
    fac = serve.RequestFactory(nmin=1e4, nmax=3e4, families=("slab", "bar"))
    reqs = [fac.make(r) for r in range(1, 25)]
    with engine.Server(step_bundles=60_000) as sv:
        ids = [sv.submit(*r[1:]) for r in reqs]
        gone = ids[5::6]
        for rid in gone:
            sv.cancel(rid)
        keep = [rid for rid in ids if rid not in gone]
        res = drain(sv, keep)
        assert sorted(res) == sorted(keep)
        for rid in gone:
            with pytest.raises(engine.EngineError) as e:
                sv.poll(rid)
            assert e.value.code == abi.MCG_ENOTFOUND
    for rid, r in zip(ids, reqs):
        if rid in gone:
            continue
        with engine.Scene(r[1], r[2]) as sc:
            f, st, _ = sc.solve(r[3])
        assert res[rid][1]["steps"] == st["steps"]
        assert np.all(rownorm_rel(res[rid][0], f) < 1e-10)
