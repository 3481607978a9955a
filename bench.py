"""Benchmark of the phonon trajectory engine on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the 3-D bar the headline metric is quoted on): a 1 um x 100 nm x
100 nm bar of the Si-like three-branch material at T_ref = 300 K, ends
isothermal at +10 K / -10 K, four diffuse side walls, 1e7 energy bundles,
100 x 1 x 1 Multi tally (temperature + heat flux), FP32 flight with FP64
tallies and sampling tables.  One timed step is one complete field solve
(mcg_solve) of all 1e7 bundles with the scene already resident on the GPU.

Printed: one JSON line with the scatter-step throughput, the roofline
object of the trajectory kernel (algorithmic bytes = 96 B of bundle state
round trip per scatter-step), the tally error of the FP32-flight run against
the FP64 run of the same streams, the FP64 CPU checker timed on the same
workload on the host cores (whose result for the first seed the GPU's FP64
and FP32 runs are also compared with), and secondary objects:
"bar_streamed" (the same K bar solves submitted together as one
mcg_batch_solve, so that each solve's drain is filled by the next), "slab" (the 1-D slab of 1e5 bundles that the reference runs on its CPU, in
deterministic FP64, beside the whole solve on the CPU checker), "porous_film"
(the periodic porous film with a cylindrical pore, 1e8 bundles, 64 x 64 map,
with its own CPU baseline sample, FP32-vs-FP64 errors and the FP64 rate),
"octet" (the
reference's 42-cell octet-truss unit, FP32 and FP64, with a CPU sample),
"batched" (256
independent slab solves in one persistent launch, its graph replay and the
same solves served by the continuous-batching server, against the same
solves run one mcg_solve at a time), "serving" (open-loop Poisson request
streams through the server at six offered loads, with p50/p99 latency, the
same mix as one batch and one solve at a time, a saturating stream of small
requests, and FP64 re-solves of 1 % of the requests on the GPU and the CPU
checker) and "engines" (the history kernel against the event engine on the
bar, the porous film and the octet truss).
--headline-only skips the secondary objects and the CPU leg.

    python bench.py --only bar_streamed|slab|porous_film|octet|batched|serving|engines

runs one secondary object alone (for profiling it).

    python bench.py --copy-probe

runs only a 1 GiB device-to-device copy (torch): profiling it with
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE gives the counters' scale for a
known traffic, against which the trajectory kernel's counters are read.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from montecarlocpp_amd import abi, engine, materials, scenes  # noqa: E402

METRIC = "phonon scatter-steps/s at ≤1% tally error; p50/p99 solve latency (serve)"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_STEP_FP32 = 96   # 48 B SoA bundle state read + written per scatter-step
BYTES_PER_STEP_FP64 = 152  # 76 B state in FP64
NBUNDLES = 10_000_000
CPU_REPS = 3
KERNEL_NAME = "trajectory_kernel<float, true, true, 0>"   # <F, tables in LDS, tally in LDS, walk axes>


def workload():
    mat = materials.si_like(300.0)
    dom = scenes.bar(1e-6, 1e-7, 1e-7, 100, t_hot=10.0, t_cold=-10.0)
    return mat, dom


def progress(msg):
    """One line per finished phase on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()


def cpu_median(solves, reps=CPU_REPS, fields=None):
    """The FP64 checker over (material, domain, problem) solves, reps times:
    the median rate over the reps, each rep's time being the checker's own
    solve timer (steady_clock around the solve, tables and set-up excluded)
    summed over the solves; the checker runs OpenMP static over bundles.
    A list passed as `fields` receives the last rep's results."""
    from oracle import oracle  # CPU baseline leg only
    rates, steps, secs = [], 0, []
    for _ in range(reps):
        t = 0.0
        steps = 0
        out = []
        for mat, dom, prob in solves:
            f, cst, _ = oracle.solve(mat, dom, prob, threads=cpu_threads())
            t += cst["kernel_ms"] * 1e-3
            steps += cst["steps"]
            out.append(f)
        secs.append(t)
        rates.append(steps / t)
    if fields is not None:
        fields[:] = out
    return float(np.median(rates)), int(steps), [round(x, 3) for x in secs]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the bar workload (no CPU leg, batched or serving objects)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy for counter calibration and exit")
    ap.add_argument("--traffic-json",
                    default=os.path.join(ROOT, "profiles", "trajectory_kernel_bar1e7_fp32_traffic.json"),
                    help="counter-measured HBM bytes per launch of the kernel (rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE, scaled by the copy-probe calibration)")
    ap.add_argument("--quick", action="store_true",
                    help="with --only engines: one launch shape per engine, one timed solve each")
    ap.add_argument("--only", choices=("bar_streamed", "slab", "porous_film", "octet", "batched", "serving",
                                       "engines"),
                    help="run one secondary workload alone and print its JSON object")
    args = ap.parse_args()
    if args.only:
        run = {"bar_streamed": lambda: bar_streamed_alone(args.steps),
               "slab": lambda: bench_slab(cpu=not args.no_cpu),
               "porous_film": lambda: bench_porous(cpu=not args.no_cpu),
               "octet": lambda: bench_octet(cpu=not args.no_cpu),
               "batched": lambda: bench_batched(cpu=not args.no_cpu),
               "serving": lambda: bench_serving(cpu=not args.no_cpu),
               "engines": lambda: bench_engines(quick=args.quick)}[args.only]
        print(json.dumps({args.only: run()}))
        return
    if args.copy_probe:
        import torch
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": x.numel() * 8, "copies": 3}))
        return
    if args.headline_only:
        args.no_cpu = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); a solve would shard by "
              "bundle range across GPUs, which is not implemented.", file=sys.stderr)
        sys.exit(2)

    mat, dom = workload()
    seed = 1
    with engine.Scene(mat, dom) as sc:
        prob = scenes.problem(NBUNDLES, tally=abi.MULTI, seed=seed, precision=abi.FP32_FLIGHT)
        for w in range(args.warmup):
            sc.solve(scenes.problem(NBUNDLES, seed=1000 + w, precision=abi.FP32_FLIGHT))
        steps = 0
        kern_ms = []
        fields = []
        lost = 0
        t0 = time.perf_counter()
        for k in range(args.steps):
            f, st, _ = sc.solve(scenes.problem(NBUNDLES, seed=seed + k, precision=abi.FP32_FLIGHT))
            steps += st["steps"]
            lost += st["escapes"]
            kern_ms.append(st["kernel_ms"])
            fields.append(f)
        elapsed = time.perf_counter() - t0
        progress(f"headline: {args.steps} solves in {elapsed:.2f} s")
        # the reference's precision: the same K solves in FP64 (the
        # deterministic mode), timed the same way; the first is the
        # precision split's comparison (same streams as the first FP32 solve)
        sc.solve(scenes.problem(NBUNDLES, seed=999, precision=abi.FP64))
        steps64 = 0
        f64 = None
        t0 = time.perf_counter()
        for k in range(args.steps):
            f, st64, _ = sc.solve(scenes.problem(NBUNDLES, seed=seed + k, precision=abi.FP64))
            steps64 += st64["steps"]
            if f64 is None:
                f64 = f
        elapsed64 = time.perf_counter() - t0
        progress(f"FP64: {args.steps} solves in {elapsed64:.2f} s")
        streamed = None if args.headline_only else bench_bar_streamed(sc, seed, args.steps, fields)

    value = steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    mean_kern_s = float(np.mean(kern_ms)) * 1e-3
    steps_per_launch = steps / args.steps
    achieved = steps_per_launch * BYTES_PER_STEP_FP32 / mean_kern_s

    # tally error: FP32 flight against FP64 on the same seed, and the
    # statistical standard error of the K independent solves
    f32 = fields[0]
    q = np.array([f[1].mean() for f in fields])
    Tprof = np.array([f[0] for f in fields])
    tally_error = {
        "heat_flux_fp32_vs_fp64": abs(f32[1].mean() - f64[1].mean()) / abs(f64[1].mean()),
        "temperature_fp32_vs_fp64": rel(f32[0], f64[0]),
        "heat_flux_stat_rel_se": float(q.std(ddof=1) / np.sqrt(len(q)) / abs(q.mean())),
        "temperature_stat_rel_se": float(np.linalg.norm(Tprof.std(axis=0, ddof=1))
                                         / np.sqrt(len(q)) / np.linalg.norm(Tprof.mean(axis=0))),
        "k_eff_W_per_mK": float(q.mean() * 1e-6 / 20.0),
        "k_eff_fp32_vs_fp64": abs(f32[1].mean() - f64[1].mean()) / abs(f64[1].mean()),
    }

    cpu = None
    if not args.no_cpu:
        fcpu = []
        rate, csteps, secs = cpu_median([(mat, dom, scenes.problem(NBUNDLES, seed=seed, precision=abi.FP64))],
                                        fields=fcpu)
        cpu = {"value": rate, "unit": "scatter-steps/s", "cores": cpu_threads(), "kind": "port",
               "sample": f"the whole {NBUNDLES}-bundle bar solve (seed {seed}), FP64, OpenMP static schedule "
                         f"over bundles, median of {CPU_REPS} (solve times {secs} s, {csteps} scatter-steps each)"}
        # the same solve (seed 1) against the CPU checker's FP64 result: the
        # GPU's FP64 run (bundle for bundle the same arithmetic) and the FP32
        # flight run of the same streams
        fc = fcpu[0]
        qc = fc[1].mean()
        tally_error["vs_cpu_oracle_fp64_seed1"] = {
            "gpu_fp64": {"temperature": rel(f64[0], fc[0]), "heat_flux": abs(f64[1].mean() - qc) / abs(qc),
                         "k_eff": abs(f64[1].mean() - qc) / abs(qc)},
            "gpu_fp32_flight": {"temperature": rel(f32[0], fc[0]), "heat_flux": abs(f32[1].mean() - qc) / abs(qc),
                                "k_eff": abs(f32[1].mean() - qc) / abs(qc)},
        }
        progress(f"CPU baseline done ({secs} s)")

    traffic = None
    if args.traffic_json and os.path.exists(args.traffic_json):
        with open(args.traffic_json) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    batched = serving = porous = slab = octet_ = engines = None
    if not args.headline_only:
        slab = bench_slab(cpu=not args.no_cpu)
        progress("slab done")
        porous = bench_porous(cpu=not args.no_cpu)
        progress("porous film done")
        octet_ = bench_octet(cpu=not args.no_cpu)
        progress("octet done")
        batched = bench_batched(cpu=not args.no_cpu)
        progress("batched done")
        serving = bench_serving(cpu=not args.no_cpu)
        progress("serving done")
        engines = bench_engines()
        progress("engines done")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "scatter-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "bar 1um x 100nm x 100nm, ends +/-10 K, diffuse sides, Si-like "
                        "3-branch material at 300 K, 1e7 bundles, 100x1x1 Multi tally, "
                        "FP32 flight + FP64 tallies (one step = one full solve)",
            "bundles": NBUNDLES,
            "scatter_steps_per_solve": steps_per_launch,
            "escapes": lost,
            "escape_fraction": lost / (NBUNDLES * args.steps),
        },
        "fp64": {"scatter_steps_per_s": steps64 / elapsed64, "ms_per_step": elapsed64 / args.steps * 1e3,
                 "steps": args.steps, "note": "the same K solves (seeds) in FP64, the deterministic mode"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "kernel": KERNEL_NAME,
            "kernel_ms_mean": mean_kern_s * 1e3,
            "bytes_per_step": BYTES_PER_STEP_FP32,
        },
        "tally_error": tally_error,
        "cpu_baseline": cpu,
        "bar_streamed": streamed,
        "slab": slab,
        "porous_film": porous,
        "octet": octet_,
        "batched": batched,
        "serving": serving,
        "engines": engines,
    }
    print(json.dumps(line))


def bench_bar_streamed(sc, seed, k, fields):
    """The headline's K bar solves (the same seeds) submitted together as one
    mcg_batch_solve: one launch runs them in order, so each solve's drain (its
    longest-lived bundles, one lane each, while the GPU empties) is filled
    by the next solve's bundles and only the last solve drains.  Not the
    headline value, which is solves one at a time."""
    probs = [scenes.problem(NBUNDLES, seed=seed + i, precision=abi.FP32_FLIGHT) for i in range(k)]
    engine.batch_solve([sc] * k, [scenes.problem(NBUNDLES, seed=2000 + i, precision=abi.FP32_FLIGHT)
                                  for i in range(k)])
    t0 = time.perf_counter()
    fs, st = engine.batch_solve([sc] * k, probs)
    wall = time.perf_counter() - t0
    steps = sum(x["steps"] for x in st)
    kern = st[0]["kernel_ms"] * 1e-3
    progress(f"bar streamed: {k} solves in one launch, {wall:.3f} s")
    return {"solves": k, "scatter_steps_per_s": steps / wall, "ms_per_solve": wall / k * 1e3,
            "kernel_steps_per_s": steps / kern if kern > 0 else None,
            "roofline_frac_96B": steps * BYTES_PER_STEP_FP32 / wall / HBM_PEAK,
            "escapes": sum(x["escapes"] for x in st),
            "vs_one_at_a_time_rel_diff": max(rel(a, b) for a, b in zip(fs, fields)),
            "note": "the headline's K solves (same seeds) as one mcg_batch_solve; wall clock "
                    "including the batch's set-up and result copies"}


def bar_streamed_alone(k):
    """bench_bar_streamed by itself (for profiling it): the K one-at-a-time
    solves it is compared with, then the K solves as one batch launch."""
    mat, dom = workload()
    with engine.Scene(mat, dom) as sc:
        fields = [sc.solve(scenes.problem(NBUNDLES, seed=1 + i, precision=abi.FP32_FLIGHT))[0] for i in range(k)]
        return bench_bar_streamed(sc, 1, k, fields)


def bench_porous(cpu=True, nbundles=100_000_000, cpu_sample=1_000_000):
    """Periodic porous film (cylindrical pore, 64 x 64 map), 1e8 bundles,
    FP32 flight, maxscat 100: one full solve per timed step."""
    from montecarlocpp_amd import workloads
    mat, dom, prob = workloads.porous_film(nbundles)
    vol = engine.cell_volumes(dom)
    with engine.Scene(mat, dom) as sc:
        # warm-up, kept for the precision split: the same 1e7-bundle streams
        # in FP64 (the deterministic mode, equal to the FP64 oracle)
        f_warm32, _, _ = sc.solve(workloads.porous_film(nbundles // 10)[2])
        f_warm64, st64, _ = sc.solve(workloads.porous_film(nbundles // 10, precision=abi.FP64)[2])
        # the statistical noise of a 1e7-bundle solve, for scale: FP32, another seed
        f_seed2, _, _ = sc.solve(workloads.porous_film(nbundles // 10, seed=2)[2])
        runs = []
        for k in range(2):
            p = workloads.porous_film(nbundles, seed=1 + k)[2]
            t0 = time.perf_counter()
            f, st, _ = sc.solve(p)
            runs.append((time.perf_counter() - t0, f, st))
            progress(f"porous film solve {k}: {st['steps']} steps in {runs[-1][0]:.2f} s")
    wall = sum(r[0] for r in runs)
    steps = sum(r[2]["steps"] for r in runs)
    kern = sum(r[2]["kernel_ms"] for r in runs) * 1e-3
    keff = [workloads.k_eff_porous(r[1], vol) for r in runs]
    out = {"workload": "porous film 300x300x100 nm, centred pore r = 100 nm (diffuse), periodic "
                       "x (dT 1 K) and y, diffuse top/bottom, Si-like at 300 K, 1e8 bundles, "
                       "64x64 Multi map, FP32 flight, maxscat 100",
           "bundles": nbundles, "scatter_steps_per_solve": steps / len(runs),
           "escapes": int(sum(r[2]["escapes"] for r in runs)),
           "steps_per_s": steps / wall, "kernel_steps_per_s": steps / kern,
           "ms_per_solve": wall / len(runs) * 1e3,
           "k_eff_W_per_mK": float(np.mean(keff)), "k_eff_seed_spread": float(abs(keff[0] - keff[1])),
           # FP32 flight against FP64 (same streams, 1e7 bundles); the diffuse
           # walls decorrelate the two after a few bounces, so the difference
           # is set by the solve's noise, shown beside it (seed 1 vs seed 2)
           "precision_error_1e7_bundles": {
               "k_eff_fp32_vs_fp64": abs(workloads.k_eff_porous(f_warm32, vol) / workloads.k_eff_porous(f_warm64, vol)
                                         - 1.0),
               "flux_x_map_fp32_vs_fp64": rel(f_warm32[1], f_warm64[1]),
               "temperature_map_fp32_vs_fp64": rel(f_warm32[0], f_warm64[0]),
               "k_eff_seed1_vs_seed2": abs(workloads.k_eff_porous(f_warm32, vol) / workloads.k_eff_porous(f_seed2, vol)
                                           - 1.0),
               "flux_x_map_seed1_vs_seed2": rel(f_warm32[1], f_seed2[1]),
               "temperature_map_seed1_vs_seed2": rel(f_warm32[0], f_seed2[0])},
           "roofline_frac_96B": steps / kern * BYTES_PER_STEP_FP32 / HBM_PEAK,
           # the deterministic mode's rate on the same film: the 1e7-bundle
           # FP64 solve above (one launch, its HIP-event time)
           "fp64_1e7_bundles": {"kernel_ms": st64["kernel_ms"], "scatter_steps": st64["steps"],
                                "kernel_steps_per_s": st64["steps"] / (st64["kernel_ms"] * 1e-3)}}
    if cpu:
        cp = workloads.porous_film(cpu_sample, precision=abi.FP64)[2]
        rate, csteps, secs = cpu_median([(mat, dom, cp)])
        out["cpu_baseline"] = {"value": rate, "unit": "scatter-steps/s", "cores": cpu_threads(), "kind": "port",
                               "sample": f"a {cpu_sample}-bundle solve of the same porous film, FP64, median of "
                                         f"{CPU_REPS} (solve times {secs} s, {csteps} scatter-steps each)"}
    return out


def bench_slab(cpu=True, reps=20):
    """The reference's own CPU-runnable case: the 1-D cross-plane slab
    (100 nm, walls at +/-1 K, 50 bins, 1e5 bundles, seed 1), deterministic FP64
    on the GPU (bundle for bundle the oracle's result) and, for comparison,
    the whole solve on the CPU oracle."""
    from montecarlocpp_amd import workloads
    mat, dom, prob = workloads.slab()
    with engine.Scene(mat, dom) as sc:
        sc.solve(prob)
        t0 = time.perf_counter()
        for _ in range(reps):
            f, st, _ = sc.solve(prob)
        wall = (time.perf_counter() - t0) / reps
    out = {"workload": "1-D slab 100 nm, isothermal walls +/-1 K, specular sides, Si-like at 300 K, "
                       "1e5 bundles, 50 bins, FP64 deterministic mode",
           "bundles": st["bundles"], "scatter_steps": st["steps"],
           "ms_per_solve": wall * 1e3, "kernel_ms": st["kernel_ms"],
           "steps_per_s": st["steps"] / wall}
    if cpu:
        from oracle import oracle  # CPU baseline leg only
        rate, csteps, secs = cpu_median([(mat, dom, prob)])
        fo, _, _ = oracle.solve(mat, dom, prob, threads=cpu_threads())
        out["cpu_baseline"] = {"value": rate, "unit": "scatter-steps/s", "cores": cpu_threads(), "kind": "port",
                               "sample": f"the whole 1e5-bundle solve, FP64, median of {CPU_REPS}",
                               "ms_per_solve": float(np.median(secs)) * 1e3}
        out["gpu_vs_cpu_tally_rel_diff"] = rel(f, fo)
    return out


def bench_octet(cpu=True, nbundles=10_000_000, cpu_sample=200_000):
    """The reference's octet-truss unit (OctetDomain, 42 cells: 4-gon prisms
    and pyramids, triangular prisms, boxes; 71 interface and 5 periodic
    pairs), the struts' gradient source, maxscat 100: FP32 flight and the
    deterministic FP64 mode on the GPU, and the CPU oracle on a sample."""
    from montecarlocpp_amd import materials, octet
    mat = materials.si_like(300.0)
    dom = octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0)
    out = {"workload": "octet truss unit s = 600 nm, d1 = d2 = 100 nm, t = 10 nm, div (4, 2, 2, 1), "
                       "dT 1 K, Si-like at 300 K, 1e7 bundles, Multi tally, maxscat 100"}
    with engine.Scene(mat, dom) as sc:
        for prec, key in ((abi.FP32_FLIGHT, "fp32"), (abi.FP64, "fp64")):
            prob = scenes.problem(nbundles, maxscat=100, seed=1, precision=prec)
            sc.solve(scenes.problem(nbundles // 10, maxscat=100, seed=2, precision=prec))
            f, st, _ = sc.solve(prob)
            out[key] = {"scatter_steps": st["steps"], "kernel_ms": st["kernel_ms"],
                        "steps_per_s": st["steps"] / (st["kernel_ms"] * 1e-3), "escapes": st["escapes"]}
    if cpu:
        prob = scenes.problem(nbundles, maxscat=100, seed=1, count=cpu_sample)
        rate, csteps, secs = cpu_median([(mat, dom, prob)])
        out["cpu_baseline"] = {"value": rate, "unit": "scatter-steps/s", "cores": cpu_threads(), "kind": "port",
                               "sample": f"first {cpu_sample} of the 1e7 bundles, FP64, median of {CPU_REPS} "
                                         f"(solve times {secs} s)"}
    return out


def served_wait(sv, ids):
    """Steps the server until every id has a result; returns {id: (field, stats)}."""
    want, res = set(ids), {}
    while len(res) < len(want):
        sv.step()
        for rid in sv.completed():
            if rid in want:
                res[rid] = sv.poll(rid)
    return res


def bench_batched(cpu=True):
    """256 independent slabs (random thickness, T_ref, lateral specularity,
    bundle count): one persistent launch with the autotuned batch shape, its
    graph replay, and one mcg_solve per slab; beside the whole batch on the
    FP64 checker (one solve per slab, median of 3)."""
    from montecarlocpp_amd import workloads
    cases = workloads.slab_batch(256, precision=abi.FP32_FLIGHT)
    scs = [engine.Scene(m, d) for m, d, _ in cases]
    probs = [p for _, _, p in cases]
    tuned, tuned_rate = engine.autotune_batch(scs, probs, sample_fraction=1.0)
    engine.batch_solve(scs, probs, tuned)
    t0 = time.perf_counter()
    _, st = engine.batch_solve(scs, probs, tuned)
    tb = time.perf_counter() - t0
    steps = sum(x["steps"] for x in st)
    lost = sum(x["escapes"] for x in st)
    # the default shape, for comparison
    engine.batch_solve(scs, probs)
    t0 = time.perf_counter()
    _, std = engine.batch_solve(scs, probs)
    td = time.perf_counter() - t0
    # the same batch captured once into a HIP graph and replayed
    with engine.Batch(scs, probs, tuned) as gb:
        gb.run()
        t0 = time.perf_counter()
        _, gst = gb.run()
        tg = time.perf_counter() - t0
    gsteps = sum(x["steps"] for x in gst)
    t0 = time.perf_counter()
    single_kernel_ms = 0.0
    for sc, p in zip(scs, probs):
        single_kernel_ms += sc.solve(p)[1]["kernel_ms"]
    ts = time.perf_counter() - t0
    for sc in scs:
        sc.close()
    # the same 256 solves submitted at once to the continuous-batching
    # server: from the first submission to the last result (material tables
    # and uploads included)
    with engine.Server() as sv:
        warm = [sv.submit(*cases[i]) for i in range(4)]
        served_wait(sv, warm)
        t0 = time.perf_counter()
        ids = [sv.submit(m, d, p) for m, d, p in cases]
        res = served_wait(sv, ids)
        tv = time.perf_counter() - t0
    served_steps = sum(r[1]["steps"] for r in res.values())
    out = {"workload": "256 slabs, L log-uniform 10 nm-10 um, T_ref 100-500 K, dT 2 K, "
                       "lateral specularity U[0,1], 1e4-1e6 bundles each (5.4e7 total), FP32 flight",
           "bundles": int(sum(p.nemit for p in probs)), "scatter_steps": int(steps), "escapes": int(lost),
           "autotuned_shape": {"block": tuned.block, "chunk": tuned.chunk, "lds_tables": tuned.lds_tables,
                               "sample_steps_per_s": tuned_rate},
           "batched_steps_per_s": steps / tb, "batched_ms": tb * 1e3,
           "batched_kernel_ms": st[0]["kernel_ms"],
           "default_shape_steps_per_s": sum(x["steps"] for x in std) / td,
           "graph_replay_steps_per_s": gsteps / tg, "graph_replay_ms": tg * 1e3,
           "single_solve_path_steps_per_s": steps / ts, "single_solve_path_ms": ts * 1e3,
           "single_solve_path_kernel_ms": single_kernel_ms,
           "speedup_vs_single_solve_path": ts / tb,
           "served_steps_per_s": served_steps / tv, "served_ms": tv * 1e3,
           "served_vs_single_solve_path": (served_steps / tv) / (steps / ts)}
    if cpu:
        c64 = workloads.slab_batch(256, precision=abi.FP64)
        rate, csteps, secs = cpu_median(c64)
        out["cpu_baseline"] = {"value": rate, "unit": "scatter-steps/s", "cores": cpu_threads(), "kind": "port",
                               "sample": f"all 256 slab solves in FP64, one checker solve each, median of "
                                         f"{CPU_REPS} (batch times {secs} s, {csteps} scatter-steps)"}
    return out


def bench_engines(reps=2, quick=False):
    """History kernel against the event engine (structure-of-arrays state in
    HBM, bundles regrouped by next event and cell kind each round) on the bar,
    the porous film and the octet truss, FP32 flight: the best of `reps`
    kernel times per launch shape, over a sweep of the event engine's knobs
    (bundles in flight, bundles per thread, rounds between regroupings)."""
    from montecarlocpp_amd import octet, workloads
    cases = {
        "bar_1e7": (materials.si_like(300.0), scenes.bar(1e-6, 1e-7, 1e-7, 100, t_hot=10.0, t_cold=-10.0),
                    scenes.problem(10_000_000, seed=1, precision=abi.FP32_FLIGHT)),
        "porous_2e7": workloads.porous_film(20_000_000),
        "octet_1e7": (materials.si_like(300.0), octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0),
                      scenes.problem(10_000_000, maxscat=100, seed=1, precision=abi.FP32_FLIGHT)),
    }
    shapes = [(1 << 20, 1, 1), (1 << 21, 1, 1), (1 << 22, 1, 1), (1 << 21, 4, 1), (1 << 21, 1, 4), (1 << 21, 4, 4)]
    if quick:   # one shape per engine, one timed solve each (for profiling)
        shapes, reps = [(1 << 22, 1, 1)], 1
    out = {}
    for name, (mat, dom, prob) in cases.items():
        row = {}
        with engine.Scene(mat, dom) as sc:
            sc.solve(prob)
            best = None
            for _ in range(reps):
                _, st, _ = sc.solve(prob)
                best = st if best is None or st["kernel_ms"] < best["kernel_ms"] else best
            row["history"] = {"kernel_ms": best["kernel_ms"], "steps_per_s": best["steps"] / (best["kernel_ms"] * 1e-3)}
            ev = []
            for live, pt, ce in shapes:
                t = abi.Tune()
                t.engine, t.live, t.per_thread, t.compact_every = abi.ENGINE_EVENT, live, pt, ce
                sc.set_tune(t)
                sc.solve(prob)
                best = None
                for _ in range(reps):
                    _, st, _ = sc.solve(prob)
                    best = st if best is None or st["kernel_ms"] < best["kernel_ms"] else best
                ev.append({"live": live, "per_thread": pt, "compact_every": ce, "kernel_ms": best["kernel_ms"],
                           "steps_per_s": best["steps"] / (best["kernel_ms"] * 1e-3)})
            sc.set_tune(None)
        row["event"] = ev
        row["event_best_vs_history"] = max(e["steps_per_s"] for e in ev) / row["history"]["steps_per_s"]
        out[name] = row
        progress(f"engines {name}: history {row['history']['steps_per_s']:.3g}, "
                 f"event best {max(e['steps_per_s'] for e in ev):.3g} steps/s")
    return out


def bench_serving(lams=(10, 25, 50, 100, 200, 500), nreq=100, cpu=True, cpu_cap=200_000):
    """Open-loop Poisson streams of slab / bar / random-slab / porous-film
    requests (1e4-1e8 bundles, T_ref 100-500 K) through the
    continuous-batching server, one stream of nreq requests per offered load
    lambda.  1 % of each stream (at least one request) is checked: its served
    FP32-flight result against the GPU's FP64 rerun of the whole request (same
    streams: the precision error), and its first cpu_cap bundles solved in
    FP64 by the CPU checker and by the GPU (the same bundles: agreement of the
    GPU's FP64 path with the checker); the error against a full FP64 CPU
    rerun is the first plus the second."""
    from montecarlocpp_amd import serve
    fac = serve.RequestFactory()
    small = serve.RequestFactory(nmax=1e6)
    sweep = []
    checks = []

    def same_requests(factory, n):
        """The requests of one stream (the ids of every stream are 1..n)
        as one mcg_batch_solve, and one mcg_solve each in turn (the plain
        single-solve path): steps/s of each."""
        reqs = [factory.make(r + 1) for r in range(n)]
        scs = [engine.Scene(m, d) for _, m, d, _ in reqs]
        probs = [p for _, _, _, p in reqs]
        t0 = time.perf_counter()
        _, bst = engine.batch_solve(scs, probs)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        single_steps = sum(sc.solve(p)[1]["steps"] for sc, p in zip(scs, probs))
        ts = time.perf_counter() - t0
        for sc in scs:
            sc.close()
        return sum(x["steps"] for x in bst) / tb, single_steps / ts

    with engine.Server() as sv:
        for lam in lams:
            r = serve.run(lam, nreq, server=sv, factory=fac, seed=int(lam), rerun_fraction=0.01,
                          keep_sampled=True)
            sampled = r.pop("_sampled", [])
            sweep.append(r)
            progress(f"serving lambda {lam}: {r['solves_per_s']:.1f} solves/s, "
                     f"p50 {r['latency_p50_s'] * 1e3:.1f} ms, p99 {r['latency_p99_s']:.2f} s")
            for (fam, mat, dom, prob, served), err in zip(sampled, r["fp64_rerun_errors"]):
                checks.append((lam, fam, mat, dom, prob, err))
        # a saturating stream of small requests only (1e4-1e6 bundles, all
        # arriving within a millisecond), where one solve cannot fill the GPU
        rs = serve.run(1e6, 2 * nreq, server=sv, factory=small, seed=7)
    mix_rate, single_rate = same_requests(fac, nreq)
    small_batch, small_single = same_requests(small, 2 * nreq)
    saturated = max(r["scatter_steps_per_s"] for r in sweep)
    out = {"requests_mix": "slab / bar / random slab / porous film, bundles log-uniform 1e4-1e8, "
                           "T_ref 100-500 K, FP32 flight; one server, streams run back to back",
           "requests_per_lambda": nreq, "sweep": sweep,
           "same_requests_one_batch_steps_per_s": mix_rate,
           "same_requests_single_solves_steps_per_s": single_rate,
           "saturated_served_vs_batch": saturated / mix_rate,
           "saturated_served_vs_single_solve_path": saturated / single_rate,
           "small_requests": {"mix": "the same families, bundles log-uniform 1e4-1e6, all arriving at once",
                              "served": rs, "same_requests_one_batch_steps_per_s": small_batch,
                              "same_requests_single_solves_steps_per_s": small_single,
                              "served_vs_single_solve_path": rs["scatter_steps_per_s"] / small_single,
                              "batch_vs_single_solve_path": small_batch / small_single}}
    progress(f"serving small requests: {rs['scatter_steps_per_s']:.3g} steps/s served, "
             f"{small_single:.3g} one solve at a time")
    if cpu:
        from oracle import oracle  # checker leg, after the timed streams
        rows = []
        for lam, fam, mat, dom, prob, err in checks:
            n = min(int(prob.nemit), cpu_cap)
            p64 = scenes.problem(prob.nemit, tally=prob.tally, maxscat=prob.maxscat, maxloop=prob.maxloop,
                                 seed=prob.seed, precision=abi.FP64, size=prob.size, count=n)
            fc, _, _ = oracle.solve(mat, dom, p64)
            with engine.Scene(mat, dom) as sc:
                fg, _, _ = sc.solve(p64)
            gc = max(float(np.linalg.norm(fg[k] - fc[k]) / np.linalg.norm(fc[k])) for k in range(fc.shape[0])
                     if np.linalg.norm(fc[k]) > 0)
            rows.append({"lambda_per_s": lam, "family": fam, "bundles": int(prob.nemit), "cpu_bundles": n,
                         "served_fp32_vs_gpu_fp64": {k: err[k] for k in ("temperature", "heat_flux",
                                                                         "temperature_in_se", "heat_flux_in_se")},
                         "gpu_fp64_vs_cpu_fp64_same_bundles": gc,
                         "vs_cpu_fp64_bound": max(err["temperature"], err["heat_flux"]) + gc})
        out["cpu_fp64_checks"] = rows
    return out


if __name__ == "__main__":
    main()
