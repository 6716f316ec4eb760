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
events on the library's stream during one more FP32 solve after the timed ones, for the roofline. The CPU baseline is the
oracle's serial restatement of the reference CPU path timed on a bounded sample of iterations.

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


def read_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary, if any."""
    f = ROOT / "profiles" / "traffic_fine_chebyshev_plate606k.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(workload, gpu_iterations, sample_iterations):
    """The oracle (serial C++ restatement of the reference CPU path) on the same plate: setup once,
    then `sample_iterations` PCG iterations with the f32 V-cycle, scaled to a full solve."""
    import refcpu

    t0 = time.perf_counter()
    problem = refcpu.Problem(workload)
    mg = refcpu.Multigrid(problem, refcpu.options(precision=2))
    prep = time.perf_counter() - t0
    _, r = mg.solve(1e-8, iteration_limit=sample_iterations)
    per_iteration = r.solve_seconds / max(1, r.iterations)
    return {
        "value": per_iteration * gpu_iterations * 1e3,
        "unit": "ms/solve",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (serial C++ restatement of the reference CPU path, Precision::Single V-cycle) on the same "
                   f"606k-DOF plate: {r.iterations} PCG iterations timed ({r.solve_seconds:.2f} s, "
                   f"{per_iteration * 1e3:.1f} ms/iteration) x {gpu_iterations} iterations of the GPU solve; "
                   f"assembly+setup {prep:.1f} s not included"),
    }


def multi_rhs(ctx, sf, np, columns, steps, warmup):
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
    return {
        "workload": f"block 60x40x40 mm, 30x20x20 hex x 6 Tet10, {rhs.shape[1]} free DOFs, {columns} random-traction "
                    f"load cases (seed 7), FP32 V-cycle in FP64 PCG, rtol 1e-8",
        "columns": columns,
        "ms_per_batch": round(batch_ms, 3),
        "ms_per_load_case": round(batch_ms / columns, 3),
        "sequential_ms_per_load_case": round(seq_ms / columns, 3),
        "speedup_vs_sequential": round(seq_ms / batch_ms, 3),
        "iterations": {"min": int(min(p.iteration for p in progress)), "max": int(max(p.iteration for p in progress))},
        "columns_bit_identical_to_single_solves": bool(identical),
    }


def serving_requests(sf, np, count, rate, seed):
    """Synthetic serving traffic: `count` requests with Poisson arrivals at `rate` per second. A
    request is a bar (10:1:1), a square tube (10:1:1, walls a quarter of the side) or a plate
    (1:1:0.02) of Tet10, one kind in three each, with a log-uniform size between 5k and 600k DOFs,
    E ~ U[1, 210] GPa, nu ~ U[0.25, 0.4], clamped on a random face and loaded on another by random
    tractions N(0, 1 MPa) per component and face; 30% of requests reuse the mesh, material and
    clamp of an earlier request (its cached hierarchy) with new tractions. A tube's clamp and load
    planes are its ends or its outer walls."""
    rng = np.random.default_rng(seed)
    out, systems = [], []
    t = 0.0
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if systems and rng.random() < 0.3:
            base = systems[rng.integers(len(systems))]
        else:
            dofs = float(np.exp(rng.uniform(np.log(5e3), np.log(6e5))))
            kind = rng.integers(3)
            tube = {}
            if kind == 0:  # bar: 3 (20k + 1)(2k + 1)^2 DOFs
                k = max(1, int(round((dofs / 240.0) ** (1.0 / 3.0))))
                while k > 1 and 3 * (20 * k + 1) * (2 * k + 1) ** 2 > 6e5:
                    k -= 1
                size, div = (10.0 * k, 1.0 * k, 1.0 * k), (10 * k, k, k)
            elif kind == 1:  # tube of side k cells, k/4 through each wall: 3 (20k + 1)((2k + 1)^2 - (k - 1)^2) DOFs
                k = max(4, 4 * int(round((dofs / 180.0) ** (1.0 / 3.0) / 4.0)))
                while k > 4 and 3 * (20 * k + 1) * ((2 * k + 1) ** 2 - (k - 1) ** 2) > 6e5:
                    k -= 4
                size, div = (10.0 * k, 1.0 * k, 1.0 * k), (10 * k, k, k)
                tube = dict(wall_cells=k // 4, wall_thickness=float(k // 4))
            else:  # plate: 3 (2k + 1)^2 5 DOFs
                k = max(2, int(round(np.sqrt(dofs / 60.0))))
                while k > 2 and 15 * (2 * k + 1) ** 2 > 6e5:
                    k -= 1
                size, div = (1.0 * k, 1.0 * k, 0.4), (k, k, 2)
            clamp = int(rng.integers(6))
            base = dict(size=size, divisions=div, clamp_axis=clamp // 2,
                        clamp_value=0.0 if clamp % 2 == 0 else size[clamp // 2],
                        youngs=float(rng.uniform(1e3, 2.1e5)), poisson=float(rng.uniform(0.25, 0.4)), **tube)
            systems.append(base)
        faces = [f for f in range(6) if f != 2 * base["clamp_axis"] + (0 if base["clamp_value"] == 0.0 else 1)]
        load = faces[rng.integers(len(faces))]
        axis = load // 2
        w = sf.Workload(**base, load_axis=axis, load_value=0.0 if load % 2 == 0 else base["size"][axis],
                        traction_kind=1, traction_scale=1.0, traction_seed=seed, load_case=i)
        out.append((t, w))
    return out


def serving(sf, np, count, rate, seed):
    """Serves the synthetic traffic on one GPU and reports latency percentiles."""
    traffic = serving_requests(sf, np, count, rate, seed)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--divisions", type=str, default="100,100,2")
    ap.add_argument("--cpu-iterations", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--columns", type=int, default=32, help="load cases of the multi-RHS measurement (0 = skip)")
    ap.add_argument("--graph", type=int, default=0, help="replay each V-cycle as a captured HIP graph (1) or launch eagerly (0)")
    ap.add_argument("--serve-requests", type=int, default=512, help="requests of the serving measurement (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate of the serving measurement, per s")
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
                        f"({n_free} free), jitter 5% seed 20261202, rtol 1e-8",
            "levels": int(report.multigrid.levels),
            "coarse_solve": "explicit f64 inverse (fixed-order product)",
            "operator_complexity": round(report.multigrid.operator_complexity, 4),
        },
        "fp64_ms_per_solve": round(results["fp64"]["ms"], 3),
        "fp32_over_fp64_speedup": round(results["fp64"]["ms"] / fp32["ms"], 3),
        "iterations": {"fp32": int(fp32["iterations"]), "fp64": int(results["fp64"]["iterations"])},
        "relative_residual": {"fp32": fp32["residual"], "fp64": results["fp64"]["residual"]},
        "fp32_vs_fp64_displacement_max_rel_diff": disp_err,
        "dof_iterations_per_s": {"fp32": n_free * fp32["iterations"] / (fp32["ms"] * 1e-3),
                                 "fp64": n_free * results["fp64"]["iterations"] / (results["fp64"]["ms"] * 1e-3)},
        "solves_per_s": round(1e3 / fp32["ms"], 3),
        "build_seconds": {"assembly": round(report.assembly_seconds, 3),
                          "multigrid_setup": round(report.preconditioner_seconds, 3)},
        "roofline": {
            "kernel": "cheb_step_kernel<3, float, float> (fused Chebyshev step, finest level)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(),
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_us": round(seconds / launches * 1e6, 2),
            "launches": launches,
            "achievable_copy_GBs": round(copy_gbs, 1),
            "achievable_read_GBs": round(read_gbs, 1),
        },
        "other_kernels": {
            "system_product_f64": {"avg_launch_us": round(spmv_s / max(1, spmv_n) * 1e6, 2),
                                   "achieved_GBs": round(spmv_b / max(spmv_s, 1e-30) / 1e9, 1), "launches": spmv_n},
            "fine_residual_f32": {"avg_launch_us": round(res_s / max(1, res_n) * 1e6, 2),
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
    if args.columns:
        line["multi_rhs"] = multi_rhs(ctx, sf, np, args.columns, args.steps, args.warmup)
    if args.serve_requests:
        line["serving"] = serving(sf, np, args.serve_requests, args.serve_rate, 20261202)
        line["served_p99_ms"] = line["serving"]["p99_ms"]
    if not args.no_cpu_baseline:
        log("timing the CPU baseline (oracle) on a bounded sample")
        line["cpu_baseline"] = cpu_baseline(workload, int(fp32["iterations"]), args.cpu_iterations)
        import refcpu
        problem = refcpu.Problem(workload)
        t = time.perf_counter()
        problem.recover(nodal_u)
        line["stress_recovery"]["oracle_serial_ms"] = round((time.perf_counter() - t) * 1e3, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
