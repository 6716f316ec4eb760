#!/usr/bin/env python3
"""Benchmark of the linear-static solve path on one MI355X.

Workload: the 606,015-DOF jittered two-layer plate (100 x 100 x 0.8 mm, 100 x 100 x 2 hex cells
x 6 Tet10, E 3.5 GPa, nu 0.36, clamped at x = 0, 0.01 MPa on the top face, interior vertices
jittered by U(-5%, +5%) of the cell, seed 20261202). The matrix, the multigrid hierarchy and the
solve's vectors are assembled and set up once and stay resident in HBM; a step is one full
conjugate gradient solve from a zero estimate to a true relative residual of 1e-8.

The headline value is ms per solve with the FP32 V-cycle inside the FP64 conjugate gradient
(the reference's Precision::Single); the plain all-FP64 form (Precision::Double) is timed beside
it. The dominant kernel -- the fused Chebyshev step on the finest level -- is timed with HIP
events on the library's stream during one more FP32 solve after the timed ones, for the roofline,
and so is the conjugate gradient's FP64 product A p.

Both GPU forms are checked against the oracle's FP64 solve of the same plate in the reference's
own order (serial dot products, reference coarse solve): energy-norm and displacement errors.
The CPU baseline is the oracle (the C++ restatement of the reference CPU path, with its
row-chunk thread pool): assembly and setup timed whole, the FP32-V-cycle solve on a bounded
sample of iterations at the reference's default thread count, scaled to the GPU's iteration
count; the FP64 oracle solve, run whole for the error check, is timed at all the host's cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "ms/solve to rtol 1e-8 @606k-DOF plate, FP32 vs FP64 V-cycle; served p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` (its full name, template arguments included) from the
    committed rocprofv3 PMC summary, or None when that variant was not profiled."""
    f = ROOT / "profiles" / "traffic_plate606k.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text())["kernels"].get(kernel, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_info():
    """The host's CPU model, logical processors and this process's CPU share."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = min(share, int(os.environ["OMP_NUM_THREADS"]))
    return model, logical, max(1, share)


def oracle_leg(np, workload, gpu_iterations, sample_iterations, x_gpu):
    """The oracle on the same plate. Assembly and the multigrid setup are timed whole; the
    FP32-V-cycle solve (Precision::Single) is timed on `sample_iterations` iterations at the
    reference's default thread count, min(logical / 2, 8) (parallel.rs:43-49), and scaled to the
    GPU's iteration count; the FP64 solve (Precision::Double) in the reference's order runs whole
    at all the host's cores and is the answer the GPU forms are checked against."""
    import errors
    import refcpu

    model, logical, share = cpu_info()
    default = int(refcpu.lib().ref_default_threads())
    t0 = time.perf_counter()
    problem = refcpu.Problem(workload)
    assembly = time.perf_counter() - t0
    t0 = time.perf_counter()
    mg32 = refcpu.Multigrid(problem, refcpu.options(precision=2, threads=default))
    setup = time.perf_counter() - t0
    _, r = mg32.solve(1e-8, iteration_limit=sample_iterations)
    per_iteration = r.solve_seconds / max(1, r.iterations)
    threads32 = int(mg32.threads)
    log(f"oracle: assembly {assembly:.2f} s, setup {setup:.2f} s, FP32 V-cycle {per_iteration * 1e3:.1f} ms/iteration "
        f"on {threads32} threads")
    del mg32
    t0 = time.perf_counter()
    mg64 = refcpu.Multigrid(problem, refcpu.options(precision=0, threads=share))
    setup64 = time.perf_counter() - t0
    disp, r64 = mg64.solve(1e-8)
    log(f"oracle: FP64 solve {r64.solve_seconds:.1f} s, {r64.iterations} iterations on {mg64.threads} threads")
    fm = problem.free_map()
    x_o = disp.reshape(-1, 3)[fm != 0xFFFFFFFF].reshape(-1)
    rs, cols, vals = problem.csr()
    norm_o = errors.energy(rs, cols, vals, x_o)
    check = {"oracle_iterations": int(r64.iterations), "oracle_relative_residual": r64.relative_residual,
             "oracle_at_rounding_floor": bool(r64.at_rounding_floor)}
    for name, x in x_gpu.items():
        check[f"{name}_vs_oracle_energy_rel"] = errors.energy(rs, cols, vals, x - x_o) / norm_o
        check[f"{name}_vs_oracle_displacement_max_rel"] = float(np.abs(x - x_o).max() / np.abs(x_o).max())
    baseline = {
        "value": per_iteration * gpu_iterations * 1e3,
        "unit": "ms/solve",
        "cores": threads32,
        "kind": "port",
        "sample": (f"oracle (C++ restatement of the reference CPU path, its row-chunk pool) on the same 606k-DOF plate, "
                   f"Precision::Single: {r.iterations} PCG iterations timed ({r.solve_seconds:.2f} s, "
                   f"{per_iteration * 1e3:.1f} ms/iteration on {threads32} threads, the reference's default) x "
                   f"{gpu_iterations} iterations of the GPU solve; assembly and setup timed whole, apart"),
        "cpu_model": model,
        "logical_processors": logical,
        "cpu_share": share,
        "assembly_s": round(assembly, 3),
        "multigrid_setup_s": round(setup, 3),
        "solve_ms_per_iteration": round(per_iteration * 1e3, 3),
        "fp64_full_solve": {"threads": int(mg64.threads), "setup_s": round(setup64, 3),
                            "solve_s": round(r64.solve_seconds, 3), "iterations": int(r64.iterations)},
    }
    return baseline, check


def multi_rhs(ctx, sf, np, columns, steps, warmup, sequential=True):
    """Many load cases over one hierarchy: the 60 x 40 x 40 mm block (30 x 20 x 20 hex x 6 Tet10,
    E 70 GPa, nu 0.33, clamped z = 0) under `columns` random-traction load cases (N(0, 1 MPa) per
    component on the z = 40 faces, seed 7, case c), solved as one lockstep batch with the FP32
    V-cycle, against the same cases solved one at a time on the same hierarchy."""
    from workloads import block_config

    cases = [block_config(c) for c in range(columns)]
    pcg_options = sf.PcgOptions(tolerance=1e-8, dot_order=sf.DotOrder.TREE)
    builder = sf.StudyBuilder.new(ctx, sf.Model.from_workload(cases[0]), pcg_options,
                                  sf.MultigridOptions(precision=sf.Precision.DOUBLE))
    builder.step(2 ** 62)
    study = builder.finish()
    system = study.system()
    mg = study.multigrid().with_precision(sf.Precision.SINGLE, coarse_solve=1)
    rhs = np.stack([sf.Model.from_workload(w).load_vector() for w in cases])
    log(f"multi-rhs: block with {rhs.shape[1]} free DOFs, {columns} load cases")
    batch = sf.PcgBatch.new(ctx, system, mg, rhs, pcg_options)

    def run_batch():
        batch.restart()
        running, progress = batch.step(2 ** 62)
        if running or any(p.outcome != sf.StepOutcome.CONVERGED for p in progress):
            raise RuntimeError("a batched load case did not converge")
        return progress

    for _ in range(warmup):
        run_batch()
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        progress = run_batch()
    ctx.synchronize()
    batch_ms = (time.perf_counter() - t) / steps * 1e3
    xb = batch.estimate()
    # one more batch with the batched kernels bracketed by HIP events: their rooflines
    kernels = (sf.Kernel.BATCH_FINE_CHEBYSHEV, sf.Kernel.BATCH_SYSTEM_PRODUCT)
    for k in kernels:
        sf.profile_enable(k, True)
    sf.profile_reset()
    run_batch()
    ctx.synchronize()
    timed = {k: sf.profile_read(k) for k in kernels}
    for k in kernels:
        sf.profile_enable(k, False)

    def roofline(k, what):
        n, sec, nbytes = timed[k]
        per = nbytes / max(1, n)
        gbs = per / (sec / max(1, n)) / 1e9 if n else 0.0
        return {"kernel": what, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": per,
                "avg_launch_us": round(sec / max(1, n) * 1e6, 2), "launches": n}
    out = {
        "workload": f"block 60x40x40 mm, 30x20x20 hex x 6 Tet10, {rhs.shape[1]} free DOFs, {columns} random-traction "
                    f"load cases (seed 7), FP32 V-cycle in FP64 PCG, rtol 1e-8",
        "columns": columns,
        "ms_per_batch": round(batch_ms, 3),
        "ms_per_load_case": round(batch_ms / columns, 3),
        "iterations": {"min": int(min(p.iteration for p in progress)), "max": int(max(p.iteration for p in progress))},
        "roofline_batched_fine_chebyshev": roofline(sf.Kernel.BATCH_FINE_CHEBYSHEV,
                                                    f"batched fused Chebyshev step, finest level, f32, {columns} columns"),
        "roofline_batched_system_product": roofline(sf.Kernel.BATCH_SYSTEM_PRODUCT,
                                                    f"batched system product A P, f64, {columns} columns"),
    }
    if not sequential:
        return out
    singles = [sf.Pcg.new(ctx, system, mg, rhs[m], pcg_options) for m in range(columns)]
    for p in singles[:2]:
        p.step(2 ** 62)
    ctx.synchronize()
    t = time.perf_counter()
    iters = []
    for p in singles:
        p.restart()
        iters.append(p.step(2 ** 62).iteration)
    ctx.synchronize()
    seq_ms = (time.perf_counter() - t) * 1e3
    identical = all(np.array_equal(singles[m].estimate().view(np.uint64), xb[m].view(np.uint64)) for m in range(columns))
    log(f"multi-rhs: batch {batch_ms:.1f} ms for {columns} cases, one at a time {seq_ms:.1f} ms; bit-identical {identical}")
    out["sequential_ms_per_load_case"] = round(seq_ms / columns, 3)
    out["speedup_vs_sequential"] = round(seq_ms / batch_ms, 3)
    out["columns_bit_identical_to_single_solves"] = bool(identical)
    return out


def serving(sf, np, count, rate, seed):
    """Serves the synthetic traffic on one GPU and reports latency percentiles."""
    from workloads import serving_requests

    traffic = serving_requests(count, rate, seed)
    opts = sf.ServerOptions(hbm_budget=200 << 30, max_columns=32, slice_iterations=8, builders=6, solvers=8)
    server = sf.Server.new(0, opts)
    ids = []
    t0 = time.perf_counter()
    for at, w in traffic:
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(w))
        if len(ids) % 64 == 0:
            log(f"serving: {len(ids)} submitted, {server.stats()['completed']} answered")
    states = [server.wait(i, 600.0) for i in ids]
    wall = time.perf_counter() - t0
    st = server.stats()
    done = [s for s in states if s.state == sf.ServeState.DONE]
    dofs = [s.free_dofs for s in done]
    built = [s for s in done if not s.cache_hit]

    def quantiles(values):
        if not values:
            return None
        v = np.sort(np.asarray(values))
        return {"p50_ms": round(float(v[int(np.ceil(0.5 * len(v))) - 1]) * 1e3, 1),
                "p99_ms": round(float(v[int(np.ceil(0.99 * len(v))) - 1]) * 1e3, 1)}
    server.close()
    log(f"serving: {st['completed']} answered, {st['failed']} failed in {wall:.1f} s; p50 {st['p50_seconds'] * 1e3:.0f} ms, "
        f"p99 {st['p99_seconds'] * 1e3:.0f} ms; {st['builds']} builds, {st['cache_hits']} cache hits, "
        f"{st['evictions']} evictions")
    return {
        "workload": f"{count} requests, Poisson arrivals at {rate}/s (seed {seed}); bars, tubes and plates of Tet10, "
                    f"log-uniform 5k-600k DOFs, E U[1,210] GPa, nu U[0.25,0.4], random clamp and loaded faces, "
                    f"random tractions N(0,1 MPa); 30% reuse an earlier system; HBM budget 200 GiB; FP32 V-cycle, rtol 1e-8",
        "arrival_rate_per_s": rate,
        "p50_ms": round(st["p50_seconds"] * 1e3, 1),
        "p99_ms": round(st["p99_seconds"] * 1e3, 1),
        "mean_ms": round(st["mean_seconds"] * 1e3, 1),
        "max_ms": round(st["max_seconds"] * 1e3, 1),
        "completed": int(st["completed"]),
        "failed": int(st["failed"]),
        "builds": int(st["builds"]),
        "cache_hits": int(st["cache_hits"]),
        "evictions": int(st["evictions"]),
        "hbm_peak_reserved_gib": round(st["hbm_peak"] / 2 ** 30, 2),
        "wall_s": round(wall, 1),
        "served_dofs_per_s": round(sum(dofs) / wall, 1),
        "median_request_dofs": int(np.median(dofs)) if dofs else 0,
        "builders": opts.builders,
        "solver_streams": opts.solvers,
        "latency_parts": {
            "queued": quantiles([s.queued_seconds for s in done]),
            "build_of_new_systems": quantiles([s.build_seconds for s in built]),
            "solve": quantiles([s.solve_seconds for s in done]),
            "cache_hit_latency": quantiles([s.latency_seconds for s in done if s.cache_hit]),
            "new_system_latency": quantiles([s.latency_seconds for s in built]),
        },
    }


def profile_variants(sf, np, ctx, workload):
    """Every batch depth of the finest level's kernels and of the conjugate gradient's product,
    forced in turn, each through one FP32-V-cycle and one FP64 solve of the plate, then two batched
    solves of the 32-load-case block, then the plain 1 GiB copy that calibrates the byte counters.
    The bits never depend on the depth."""
    pcg_options = sf.PcgOptions(tolerance=1e-8, dot_order=sf.DotOrder.TREE)
    for depth in (1, 2, 4, 8):
        sf.force_variant(depth)
        builder = sf.StudyBuilder.new(ctx, sf.Model.from_workload(workload), pcg_options,
                                      sf.MultigridOptions(precision=sf.Precision.DOUBLE))
        builder.step(2 ** 62)
        study = builder.finish()
        mg64 = study.multigrid().with_precision(sf.Precision.DOUBLE, coarse_solve=1)
        for mg in (mg64.with_precision(sf.Precision.SINGLE), mg64):
            p = sf.Pcg.new(ctx, study.system(), mg, study.rhs(), pcg_options).step(2 ** 62)
            log(f"depth {depth} ({mg.precision().name}): {p.iteration} iterations")
        del study, builder
    sf.force_variant(0)
    multi_rhs(ctx, sf, np, 32, 1, 0, sequential=False)
    ctx.stream_copy(1 << 30, 10)
    log("calibration copy done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--divisions", type=str, default="100,100,2")
    ap.add_argument("--cpu-iterations", type=int, default=60, help="FP32 oracle iterations timed for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--columns", type=int, default=32, help="load cases of the multi-RHS measurement (0 = skip)")
    ap.add_argument("--graph", type=int, default=0, help="replay each V-cycle as a captured HIP graph (1) or launch eagerly (0)")
    ap.add_argument("--serve-requests", type=int, default=512, help="requests of the serving measurement (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate of the serving measurement, per s")
    ap.add_argument("--multi-rhs-only", action="store_true",
                    help="only the 32-load-case block (for profiling its batched kernels); prints its multi_rhs object")
    ap.add_argument("--profile-variants", action="store_true",
                    help="for a PMC profiler: one FP32 and one FP64 solve with every batch depth forced in turn, then "
                         "the 1 GiB calibration copy; prints no line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: there is no multi-GPU path (asked for {args.gpus}); "
              "independent solves and load cases shard one per GPU with no communication", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import solidfem_amd as sf
    from workloads import plate_config

    divisions = tuple(int(v) for v in args.divisions.split(","))
    workload = plate_config(divisions=divisions)
    ctx = sf.Context(0)
    if args.profile_variants:
        profile_variants(sf, np, ctx, workload)
        return
    if args.multi_rhs_only:
        print(json.dumps(multi_rhs(ctx, sf, np, args.columns or 32, args.steps, args.warmup, sequential=False)), flush=True)
        return

    log(f"building the plate {divisions}")
    t0 = time.perf_counter()
    pcg_options = sf.PcgOptions(tolerance=1e-8, dot_order=sf.DotOrder.TREE)
    builder = sf.StudyBuilder.new(ctx, sf.Model.from_workload(workload), pcg_options,
                                  sf.MultigridOptions(precision=sf.Precision.DOUBLE))
    builder.step(2 ** 62)
    study = builder.finish()
    build_seconds = time.perf_counter() - t0
    report = study.report()
    log(f"built in {build_seconds:.1f} s: {report.free_dofs} free DOFs, {report.nonzeros} stored values, "
        f"{report.multigrid.levels} levels, operator complexity {report.multigrid.operator_complexity:.3f}; "
        f"assembly {report.assembly_seconds:.2f} s, preconditioner {report.preconditioner_seconds:.2f} s")

    system = study.system()
    rhs = study.rhs()
    # Both forms use the explicit-inverse coarse solve (a fixed-order dense product) in place of the
    # serial back substitution; the reference-order solve is kept for the parity tests.
    mg64 = study.multigrid().with_precision(sf.Precision.DOUBLE, coarse_solve=1)
    mg32 = mg64.with_precision(sf.Precision.SINGLE)
    forms = {
        "fp64": sf.Pcg.new(ctx, system, mg64, rhs, pcg_options),
        "fp32": sf.Pcg.new(ctx, system, mg32, rhs, pcg_options),
    }

    def solve(pcg):
        pcg.restart()
        p = pcg.step(2 ** 62)
        if p.outcome != sf.StepOutcome.CONVERGED:
            raise RuntimeError(f"solve did not converge: {p}")
        return p

    if args.graph:
        mg64.set_graph(True)
        mg32.set_graph(True)

    results = {}
    for name in ("fp64", "fp32"):
        pcg = forms[name]
        for _ in range(args.warmup):
            solve(pcg)
        ctx.synchronize()
        t = time.perf_counter()
        progress = [solve(pcg) for _ in range(args.steps)]
        ctx.synchronize()
        elapsed = time.perf_counter() - t
        it = progress[-1].iteration
        assert all(p.iteration == it for p in progress), "a repeated solve took a different path"
        results[name] = dict(ms=elapsed / args.steps * 1e3, iterations=it, residual=progress[-1].relative_residual,
                             floor=progress[-1].at_rounding_floor, x=pcg.estimate())
        log(f"{name}: {results[name]['ms']:.2f} ms/solve, {it} iterations, residual {progress[-1].relative_residual:.2e}")

    # One more FP32 solve with the kernels bracketed by HIP events on the library's stream: the
    # per-launch durations behind the roofline (kept out of the timed solves above).
    for k in sf.Kernel:
        sf.profile_enable(k, True)
    sf.profile_reset()
    solve(forms["fp32"])
    ctx.synchronize()
    results["fp32"]["kernels"] = {k.name: sf.profile_read(k) for k in sf.Kernel}
    for k in sf.Kernel:
        sf.profile_enable(k, False)

    copy_bytes = 1 << 30
    copy_gbs = 2 * copy_bytes / ctx.stream_copy(copy_bytes, 10) / 1e9
    read_gbs = copy_bytes / ctx.stream_read(copy_bytes, 10) / 1e9
    log(f"plain copy of 1 GiB: {copy_gbs:.0f} GB/s (read + write); plain read: {read_gbs:.0f} GB/s")

    x64, x32 = results["fp64"]["x"], results["fp32"]["x"]
    disp_err = float(np.abs(x32 - x64).max() / np.abs(x64).max())
    n_free = report.free_dofs
    fp32 = results["fp32"]
    launches, seconds, nbytes = fp32["kernels"]["FINE_CHEBYSHEV"]
    per_launch = nbytes / launches
    achieved = per_launch / (seconds / launches) / 1e9
    spmv_n, spmv_s, spmv_b = fp32["kernels"]["SYSTEM_PRODUCT"]
    res_n, res_s, res_b = fp32["kernels"]["FINE_RESIDUAL"]
    # the variants the tuning chose, by their full kernel names (the batch depth is a template
    # argument), so the traffic read from the PMC summary is that of the timed variant
    step_depth, res_depth, spmv_depth = mg32.tuned_depth(0, 0), mg32.tuned_depth(0, 1), system.tuned_depth()
    step_kernel = f"void cheb_step_kernel<3, {step_depth}, float, float>"
    spmv_kernel = f"void spmv_kernel<3, 3, {spmv_depth}, double, double, 0>"
    spmv_per_launch = spmv_b / max(1, spmv_n)
    spmv_achieved = spmv_per_launch / (spmv_s / max(1, spmv_n)) / 1e9

    line = {
        "metric": METRIC,
        "value": round(fp32["ms"], 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp32["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle in f64 PCG",
        "data": "synthetic (seeded jittered plate)",
        "config": {
            "workload": f"plate 100x100x0.8 mm, {divisions[0]}x{divisions[1]}x{divisions[2]} hex x 6 Tet10, "
                        f"{3 * (2 * divisions[0] + 1) * (2 * divisions[1] + 1) * (2 * divisions[2] + 1)} DOFs "
                        f"({n_free} free), jitter 5% seed 20261202; to rtol 1e-8 or the reference's rounding-floor "
                        f"stop (pcg.rs:234-255), whichever comes first",
            "levels": int(report.multigrid.levels),
            "coarse_solve": "explicit f64 inverse (fixed-order product)",
            "operator_complexity": round(report.multigrid.operator_complexity, 4),
        },
        "fp64_ms_per_solve": round(results["fp64"]["ms"], 3),
        "fp32_over_fp64_speedup": round(results["fp64"]["ms"] / fp32["ms"], 3),
        "iterations": {"fp32": int(fp32["iterations"]), "fp64": int(results["fp64"]["iterations"])},
        "relative_residual": {"fp32": fp32["residual"], "fp64": results["fp64"]["residual"]},
        "at_rounding_floor": {"fp32": bool(fp32["floor"]), "fp64": bool(results["fp64"]["floor"])},
        "fp32_vs_fp64_displacement_max_rel_diff": disp_err,
        "dof_iterations_per_s": {"fp32": n_free * fp32["iterations"] / (fp32["ms"] * 1e-3),
                                 "fp64": n_free * results["fp64"]["iterations"] / (results["fp64"]["ms"] * 1e-3)},
        "solves_per_s": round(1e3 / fp32["ms"], 3),
        "build_seconds": {"assembly": round(report.assembly_seconds, 3),
                          "multigrid_setup": round(report.preconditioner_seconds, 3)},
        "roofline": {
            "kernel": f"{step_kernel} (fused Chebyshev step, finest level, batch depth {step_depth})",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(step_kernel),
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_us": round(seconds / launches * 1e6, 2),
            "launches": launches,
            "achievable_copy_GBs": round(copy_gbs, 1),
            "achievable_read_GBs": round(read_gbs, 1),
        },
        "roofline_fp64_system_product": {
            "kernel": f"{spmv_kernel} (the conjugate gradient's A p, f64, batch depth {spmv_depth})",
            "bound": "hbm",
            "achieved": round(spmv_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(spmv_achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(spmv_kernel),
            "algorithmic_bytes_per_launch": spmv_per_launch,
            "avg_launch_us": round(spmv_s / max(1, spmv_n) * 1e6, 2),
            "launches": spmv_n,
        },
        "other_kernels": {
            "fine_residual_f32": {"kernel": f"void spmv_kernel<3, 3, {res_depth}, float, float, 1>",
                                  "avg_launch_us": round(res_s / max(1, res_n) * 1e6, 2),
                                  "achieved_GBs": round(res_b / max(res_s, 1e-30) / 1e9, 1), "launches": res_n},
        },
        "cpu_baseline": None,
    }
    # stress recovery of the FP32 solution (results.rs:307-576): every element's integration-point
    # stresses, the vertex patch fits and the nodal field, on the device, host arrays in and out
    model = sf.Model.from_workload(workload)
    nodal_u = np.zeros((report.total_dofs // 3, 3))
    free_map = model.free_problem()[0]
    live = free_map != 0xFFFFFFFF
    nodal_u[live] = x32.reshape(-1, 3)[free_map[live]]
    model.recover(ctx, nodal_u)  # warm
    t = time.perf_counter()
    points, nodal = model.recover(ctx, nodal_u)
    recovery_ms = (time.perf_counter() - t) * 1e3
    line["stress_recovery"] = {"ms": round(recovery_ms, 2), "elements": int(points.shape[0]),
                               "nodes": int(nodal.shape[0]), "peak_nodal_von_mises": float(nodal[:, 6].max()),
                               "note": "host incidence lists, uploads and downloads included"}
    log(f"stress recovery of the plate: {recovery_ms:.1f} ms, peak nodal von Mises {nodal[:, 6].max():.4g} MPa")
    # the error estimate of that recovery (error_estimate.rs): 14-point integrals per element on the
    # device, the sums in element order on the host
    recovery = model.begin_recovery(ctx, nodal_u, converged=True)
    recovery.advance(2 ** 62)
    model.begin_error_estimate(recovery).advance(2 ** 62)  # warm
    t = time.perf_counter()
    estimation = model.begin_error_estimate(recovery)
    estimation.advance(2 ** 62)
    estimate = estimation.estimate()
    estimate_ms = (time.perf_counter() - t) * 1e3
    line["error_estimate"] = {"ms": round(estimate_ms, 2), "relative_error": estimate.relative_error,
                              "note": "a finished recovery in; the mesh upload and the element errors' download included"}
    log(f"error estimate of the plate: {estimate_ms:.1f} ms, eta {estimate.relative_error:.4g}")
    if args.columns:
        line["multi_rhs"] = multi_rhs(ctx, sf, np, args.columns, args.steps, args.warmup)
    if args.serve_requests:
        line["serving"] = serving(sf, np, args.serve_requests, args.serve_rate, 20261202)
        line["served_p99_ms"] = line["serving"]["p99_ms"]
    if not args.no_cpu_baseline:
        log("the oracle: CPU baseline and the FP64 answer the GPU forms are checked against")
        line["cpu_baseline"], line["oracle_check"] = oracle_leg(np, workload, int(fp32["iterations"]), args.cpu_iterations,
                                                                {"fp32": x32, "fp64": x64})
        line["fp32_vs_oracle_energy_rel"] = line["oracle_check"]["fp32_vs_oracle_energy_rel"]
        line["fp64_vs_oracle_energy_rel"] = line["oracle_check"]["fp64_vs_oracle_energy_rel"]
        import refcpu
        problem = refcpu.Problem(workload)
        t = time.perf_counter()
        problem.recover(nodal_u)
        line["stress_recovery"]["oracle_serial_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        t = time.perf_counter()
        oracle_estimate = problem.error_estimate(nodal_u)
        line["error_estimate"]["oracle_serial_ms_with_recovery"] = round((time.perf_counter() - t) * 1e3, 1)
        line["error_estimate"]["oracle_bits_equal"] = bool(
            oracle_estimate["status"] == 0
            and np.array_equal(oracle_estimate["element_errors"].view(np.uint64), estimate.element_errors.view(np.uint64))
            and np.float64(oracle_estimate["relative_error"]).view(np.uint64) == np.float64(estimate.relative_error).view(np.uint64))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
