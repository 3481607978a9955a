"""Open-loop serving of field solves on one MI355X.

Requests arrive as a Poisson stream of rate lam (solves/s).  Request id r
draws, from the Philox block (counter (r, 0, 0, 0), key (0x5EED0004, 0)) as
uniforms u0..u3: its family (u0), its bundle count N = 10^(log10 nmin + (log10 nmax -
log10 nmin) u1), its T_ref = 100 + 400 u2 K and, for a random slab, its
thickness 10^(-8 + 3 u3) m; the random slab's lateral specularity is the
first word of the block (r, 1, 0, 0).  Its seed is r.  The families are the
slab, the bar, a random slab of the batched workload and the periodic porous
film.  Requests are
submitted to an mcg_server as their arrival time passes and the server is
stepped continuously; a request's latency runs from its scheduled arrival to
the step that finished it.

    python -m montecarlocpp_amd.serve --lam 100 --requests 300
"""
import argparse
import json
import math
import time

import numpy as np

from . import abi, engine, materials, scenes, workloads

REQUEST_KEY = (0x5EED0004, 0)
FAMILIES = ("slab", "bar", "random_slab", "porous_film")


class RequestFactory:
    def __init__(self, nmin=1e4, nmax=1e8, precision=abi.FP32_FLIGHT, families=FAMILIES):
        self.nmin, self.nmax, self.precision = nmin, nmax, precision
        self.families = families
        self.base = materials.si_like(300.0)
        self._mat_cache = {}

    def material(self, T):
        key = round(T, 6)
        m = self._mat_cache.get(key)
        if m is None:
            m = self.base.at_temperature(T)
            if len(self._mat_cache) < 4096:
                self._mat_cache[key] = m
        return m

    def make(self, rid):
        w = engine.philox([rid, 0, 0, 0], list(REQUEST_KEY))
        u = [x / 2.0 ** 32 for x in w]
        fam = self.families[min(int(u[0] * len(self.families)), len(self.families) - 1)]
        lo, hi = math.log10(self.nmin), math.log10(self.nmax)
        n = int(round(10.0 ** (lo + (hi - lo) * u[1])))
        T = 100.0 + 400.0 * u[2]
        if fam == "slab":
            dom = scenes.slab(100e-9, 50)
        elif fam == "bar":
            dom = scenes.bar(1e-6, 1e-7, 1e-7, 100, 10.0, -10.0)
        else:
            L = 10.0 ** (-8.0 + 3.0 * u[3])
            bins = int(round(20 + 180 * math.log10(L / 1e-8) / 3))
            spec = engine.philox([rid, 1, 0, 0], list(REQUEST_KEY))[0] / 2.0 ** 32
            dom = scenes.slab(L, bins, side=abi.PARTIAL, specularity=spec)
        if fam == "porous_film":
            dom = scenes.porous_film(workloads.PORE_DIM, workloads.PORE_RADIUS, (64, 64, 0))
            prob = scenes.problem(n, seed=rid, precision=self.precision,
                                  maxscat=workloads.PORE_MAXSCAT)
            return fam, self.material(T), dom, prob
        prob = scenes.problem(n, seed=rid, precision=self.precision)
        return fam, self.material(T), dom, prob


def _rownorm_rel(a, b):
    out = []
    for r in range(b.shape[0]):
        nb = np.linalg.norm(b[r])
        out.append(float(np.linalg.norm(a[r] - b[r]) / nb) if nb > 0 else float(np.linalg.norm(a[r])))
    return out


RERUN_PARTS = 8


def rerun_errors(reqs, results, ids):
    """Tally error of served requests against FP64 reruns of the same
    problems (same seed and streams) through mcg_solve.  The rerun is solved
    as RERUN_PARTS bundle sub-ranges, whose sum is the FP64 solve and whose
    spread gives its standard error per cell.  Reported per request, for the
    temperature row and for the heat-flux field (the three flux rows
    together): the relative difference ||served - FP64|| / ||FP64||, and the
    difference in standard errors, ||served - FP64|| / ||SE||.  Diffuse walls
    decorrelate FP32 and FP64 flight after a few bounces, so on such
    families the difference is the solve's own noise (about one SE)."""
    errs = []
    for i in ids:
        fam, mat, dom, prob = reqs[i]
        n = int(prob.nemit)
        with engine.Scene(mat, dom) as sc:
            nemit = sc.plan(prob).nemit
            cuts = np.linspace(0, nemit, RERUN_PARTS + 1).astype(np.int64)
            parts = []
            for a, b in zip(cuts[:-1], cuts[1:]):
                if b <= a:
                    continue
                p64 = scenes.problem(prob.nemit, tally=prob.tally, maxscat=prob.maxscat, maxloop=prob.maxloop,
                                     seed=prob.seed, precision=abi.FP64, size=prob.size, first=int(a),
                                     count=int(b - a))
                parts.append(sc.solve(p64)[0])
        P = np.array(parts)
        f64 = P.sum(axis=0)
        se = np.sqrt(len(P)) * P.std(axis=0, ddof=1) if len(P) > 1 else np.zeros_like(f64)
        served = np.asarray(results[i])

        def rel(rows):
            d = np.linalg.norm(served[rows] - f64[rows])
            nb = np.linalg.norm(f64[rows])
            ns = np.linalg.norm(se[rows])
            return float(d / nb) if nb > 0 else float(d), float(d / ns) if ns > 0 else 0.0

        t_rel, t_z = rel(slice(0, 1))
        q_rel, q_z = rel(slice(1, 4))
        errs.append({"request": int(i + 1), "family": fam, "bundles": n,
                     "temperature": t_rel, "heat_flux": q_rel,
                     "temperature_in_se": t_z, "heat_flux_in_se": q_z})
    return errs


def run(lam, nreq, server=None, factory=None, seed=0, step_bundles=1 << 24, verbose=False,
        rerun_fraction=0.0, keep_sampled=False):
    """Serve nreq Poisson arrivals at rate lam; returns a metrics dict.  With
    rerun_fraction > 0, that fraction of the requests (at least one) is
    solved again in FP64 afterwards and the tally errors are reported; with
    keep_sampled the sampled requests and their served results are returned
    too, under "_sampled", as (family, material, domain, problem, result)."""
    factory = factory or RequestFactory()
    own = server is None
    server = server or engine.Server(step_bundles=step_bundles)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=nreq))
    reqs = [factory.make(r + 1) for r in range(nreq)]   # built before the clock starts
    lat = np.full(nreq, np.nan)
    steps = np.zeros(nreq, np.int64)
    sampled = set()
    if rerun_fraction > 0:
        k_rr = max(1, int(round(rerun_fraction * nreq)))
        sampled = set(int(x) for x in rng.choice(nreq, size=k_rr, replace=False))
    kept = {}
    outstanding = {}
    k = 0
    done = 0
    kernel_ms0 = server.info()["kernel_ms"]
    t0 = time.perf_counter()
    while done < nreq:
        now = time.perf_counter() - t0
        while k < nreq and arrivals[k] <= now:
            _, mat, dom, prob = reqs[k]
            outstanding[server.submit(mat, dom, prob)] = k
            k += 1
        if outstanding:
            _, _, ncomp = server.step()
            if ncomp:
                tnow = time.perf_counter() - t0
                for rid in server.completed():
                    res = server.poll(rid)
                    i = outstanding.pop(rid)
                    lat[i] = tnow - arrivals[i]
                    steps[i] = res[1]["steps"]
                    if i in sampled:
                        kept[i] = res[0]
                    done += 1
        elif k < nreq:
            time.sleep(max(0.0, min(arrivals[k] - now, 0.01)))
    elapsed = time.perf_counter() - t0
    info = server.info()
    if own:
        server.close()
    out = {
        "lambda_per_s": lam,
        "requests": nreq,
        "elapsed_s": elapsed,
        "solves_per_s": nreq / elapsed,
        "scatter_steps_per_s": float(steps.sum()) / elapsed,
        "gpu_busy_frac": (info["kernel_ms"] - kernel_ms0) * 1e-3 / elapsed,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(np.mean(lat)),
    }
    if sampled:
        out["fp64_rerun_errors"] = rerun_errors(reqs, kept, sorted(sampled))
        if keep_sampled:
            out["_sampled"] = [reqs[i] + (kept[i],) for i in sorted(sampled)]
    if verbose:
        print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lam", type=float, nargs="+", default=[10, 25, 50, 100, 200, 500])
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--nmax", type=float, default=1e8)
    args = ap.parse_args()
    fac = RequestFactory(nmax=args.nmax)
    with engine.Server() as sv:
        for lam in args.lam:
            run(lam, args.requests, server=sv, factory=fac, verbose=True)


if __name__ == "__main__":
    main()
