"""Benchmark of the MI355X pressure projection (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--precision mixed|fp64]
                    [--solver mgpcg|cg] [--no-cpu-baseline]
    python bench.py --config c4 | c5 | post      (batches, serving, the post-projection step)

A step is one whole projection (fractions, labelling, assembly, solve to the residual, velocity
update) of the configured scene, with every input resident in HBM before the timed region; the
velocities are restored from a device copy at the start of each step (the projection updates
them in place), and that copy is inside the timed region.  Prints one JSON line.

Default workload (the metric's configuration): 256^3 MAC grid, liquid = 8 seeded spheres
(r ~ U[0.05,0.15]) over a pool of depth 0.25, one seeded solid box holding 10% of the volume,
N(0,1) velocities plus gravity, FP32 multigrid-preconditioned CG inside an FP64
defect-correction loop to 1e-6 relative residual in the reference's infinity norm.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: (scene kind, n, seed, residual, default precision)
    "c1": ("drop", 64, 0, 1e-8),
    "c2": ("drop_solids", 128, 1, 1e-8),
    "c3": ("spheres_box", 256, 2, 1e-6),
}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


SYMBOL = {"jacobi_L0": "k_jacobi4<1>", "jacobi_dot_L0": "k_jacobi4_dot",
          "residual_L0": "k_residual4<1>", "smooth2_L0": "k_smooth2<float, 1>",
          "cg_update_z": "k_cg_update4<2, 3>", "post_local_L0": "k_post_local4<0, 3, 1>",
          "post_local_dot_L0": "k_post_local4<1, 3, 1>", "apply_dot_p": "k_apply_dot4p"}


def measured_traffic(kernel, config):
    """HBM bytes per dispatch of `kernel` measured by the rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    summarised in profiles/ (copy-calibrated); None when no measurement covers it."""
    sym = SYMBOL.get(kernel)
    if sym is None:
        return None
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*hbm_traffic*.md"))):
        for line in open(path):
            if line.startswith("TRAFFIC ") and sym in line and f" {config} " in line:
                kv = dict(x.split("=") for x in line.split() if "=" in x)
                return {"bytes": int(kv["read_bytes"]) + int(kv["write_bytes"]),
                        "source": os.path.relpath(path, ROOT)}
    return None


STAGES = ("Projection_solid_fluid_fractions", "Projection_build_highres_linsystem", "Projection_linsolve",
          "Projection_update_velocity", "Projection_projection")


def host_info():
    """CPU model (as the reference's UI logs it, src/ui/ui.cpp:583-589) and the oracle's compiler."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import subprocess
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cc = "gcc"
    return model, f"{cc}; -O3 -fopenmp -ffp-contract=off (oracle/Makefile)"


def oracle_stages(r):
    return dict(zip(STAGES, (round(r.t_fractions, 4), round(r.t_build, 4), round(r.t_linsolve, 4),
                             round(r.t_update, 4), round(r.t_projection, 4))))


def cpu_baseline(scene, n_iter_needed, residual, mic_iterations=None):
    """Reference-faithful CPU projection: the oracle restating macpressuresolver3 + pcg, with the
    MIC(0) factor computed and applied but its output discarded (pcg_solver.h:379-384), so the
    iterates are plain CG's, the reference's threading (fractions, assembly and update on all
    threads; labelling, CG and scatter serial), timed per stage under the reference's names.

    Bounded sample: two whole projections of the same scene stopped after 20 and 40 CG iterations
    separate the per-iteration cost from the fixed cost; the solve is extrapolated to the iteration
    count of the same CG iteration on this scene (measured by the GPU's FP64 run of it).  The CPU's
    own iteration count is measured on the reference's CPU-runnable configuration (64^3 drop,
    FP64, 1e-8), solved to completion.  A second line prices the intended algorithm (the MIC(0)
    output applied) at the iteration count the oracle's own MIC(0)-PCG needs on this scene."""
    import shiokaze_amd as S
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    runs = {}
    for k in (20, 40):
        prm = O.default_params(residual=residual, max_iterations=k, mic_mode=O.MIC_FAITHFUL, threads=threads)
        r, _, _ = O.project(scene.copy(), prm, pressure64=False)
        runs[k] = r
    per_it = max((runs[40].t_linsolve - runs[20].t_linsolve) / 20.0, 1e-9)
    fixed_solve = max(runs[20].t_linsolve - 20 * per_it, 0.0)  # matrix copy + MIC(0) factor
    st = oracle_stages(runs[40])
    st["Projection_linsolve"] = round(fixed_solve + n_iter_needed * per_it, 3)
    total = (st["Projection_solid_fluid_fractions"] + st["Projection_build_highres_linsystem"] +
             st["Projection_linsolve"] + st["Projection_update_velocity"])
    st["Projection_projection"] = round(total, 3)
    # the reference's CPU-runnable configuration, solved to completion
    c1 = S.make_scene(S.SCENE_DROP, 64, seed=0)
    r1, _, _ = O.project(c1, O.default_params(residual=1e-8, mic_mode=O.MIC_FAITHFUL, threads=threads),
                         pressure64=False)
    model, compiler = host_info()
    out = {
        "value": round(1.0 / total, 6), "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": (f"oracle, reference-faithful (MIC(0) factored and applied, output discarded: plain CG "
                   f"iterates), on the same {scene.w}x{scene.h}x{scene.d} scene: whole projections stopped "
                   f"after 20 and 40 CG iterations give {per_it * 1e3:.1f} ms per iteration and "
                   f"{fixed_solve:.2f} s of matrix copy and factor; the solve is extrapolated to "
                   f"{n_iter_needed} iterations (the same CG iteration to {residual:g}, counted by the GPU's "
                   f"FP64 run of it); fractions, assembly and update threaded on {threads} threads, labelling, "
                   f"CG and scatter serial, as in the reference"),
        "cpu_model": model, "compiler": compiler,
        "threads": {"Projection_solid_fluid_fractions": threads, "Projection_build_highres_linsystem":
                    f"{threads} (labelling serial)", "Projection_linsolve": 1, "Projection_update_velocity": threads},
        "stage_s": st, "cg_iterations": int(n_iter_needed), "cg_ms_per_iteration": round(per_it * 1e3, 3),
        "c1_64_fp64_1e-8": {"iterations": int(r1.iterations), "reresid": r1.reresid, "converged": bool(r1.converged),
                            "unknowns": int(r1.n_unknowns), "solves_per_s": round(1.0 / r1.t_projection, 4),
                            "stage_s": oracle_stages(r1),
                            "what": "the CPU's own CG iteration count, solved to completion"},
    }
    if mic_iterations:
        t_mic = (st["Projection_solid_fluid_fractions"] + st["Projection_build_highres_linsystem"] +
                 fixed_solve + mic_iterations * per_it + st["Projection_update_velocity"])
        out["intended_mic0"] = {
            "value": round(1.0 / t_mic, 5), "unit": "solves/s", "iterations": int(mic_iterations),
            "what": (f"PCG with the MIC(0) output applied (pcg_solver.h:89-223), {mic_iterations} iterations to "
                     f"{residual:g} (the oracle's own run on this scene) at the same cost per iteration")}
    return out


def kernel_roofline(stats, inst_steps, config):
    """Roofline entry of the dominant kernel from the event-timed pass (per-kernel durations on the
    library's stream and the algorithmic byte model), plus a per-kernel time-share log."""
    if not stats:
        return None
    kern = {k: v for k, v in stats.items() if v["launches"] > 0 and v["total_ms"] > 0}
    # dominant kernel = the HBM-bound kernel moving the most algorithmic bytes (the event-timed
    # pass adds ~2 us per launch, which overstates the thousands of latency-bound coarse launches)
    name, st = max(kern.items(), key=lambda kv: kv[1]["bytes"])
    avg_ms = st["total_ms"] / st["launches"]
    bytes_per_launch = st["bytes"] / st["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(avg_ms * 1e3, 2), "algorithmic_bytes_per_launch": int(bytes_per_launch)}
    roofline["time_share"] = round(st["total_ms"] / sum(v["total_ms"] for v in kern.values()), 4)
    tr = measured_traffic(name, config)
    if tr is not None:
        roofline["traffic"] = tr["bytes"]
        roofline["traffic_source"] = tr["source"]
    total_ms = sum(v["total_ms"] for v in kern.values())
    total_bytes = sum(v["bytes"] for v in kern.values())
    roofline["all_kernels_gbs"] = round(total_bytes / (total_ms * 1e-3) / 1e9, 1)
    log("[bench] kernel time share:")
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["total_ms"])[:32]:
        gbs = v["bytes"] / (v["total_ms"] * 1e-3) / 1e9 if v["total_ms"] else 0
        log(f"    {k:18s} {v['total_ms'] / inst_steps:8.3f} ms/step  {v['launches'] / inst_steps:7.1f} launches/step"
            f"  {gbs:8.1f} GB/s")
    return roofline


def explicit_mg(spec):
    """Params fields of an explicit --mg pre,post,coarse_pre,coarse_post,coarse_sweeps,omega,eta[,smoother[,
    local_sweeps]] (smoother: 1 = damped Jacobi, 2 = brick-local Gauss-Seidel on the finest level;
    local_sweeps: its sweeps per finest-level application, 2 or 3)."""
    if not spec:
        return {}
    v = spec.split(",")
    out = dict(mg_pre_sweeps=int(v[0]), mg_post_sweeps=int(v[1]), mg_coarse_pre_sweeps=int(v[2]),
               mg_coarse_post_sweeps=int(v[3]), mg_coarse_sweeps=int(v[4]), mg_omega=float(v[5]),
               inner_residual=float(v[6]))
    if len(v) > 7:
        out["mg_smoother"] = int(v[7])
    if len(v) > 8:
        out["mg_local_sweeps"] = int(v[8])
    return out


def mg_controls(r):
    """The multigrid controls a result reports (autotuned or default)."""
    return {"pre": r.mg_pre_sweeps, "post": r.mg_post_sweeps, "coarse_pre": r.mg_coarse_pre_sweeps,
            "coarse_post": r.mg_coarse_post_sweeps, "coarse_solve_sweeps": r.mg_coarse_sweeps,
            "omega": r.mg_omega, "inner_residual": r.inner_residual,
            "smoother": {1: "jacobi", 2: "brick-local gauss-seidel"}.get(r.mg_smoother, r.mg_smoother),
            "local_sweeps": r.mg_local_sweeps if r.mg_smoother == 2 else None, "tuned": r.tuned}


def batched_block(args, steps, single_rounds=1, kernel_timing=False):
    """The batched workload: 64 uneven grids projected in one batched call per step (mixed
    precision, 1e-6), against the plain single-solve FP64 GPU path over the same grids one by one.
    Returns (block dict, kernel stats, instrumented steps)."""
    import shiokaze_amd as S
    from shiokaze_amd import loadgen
    from shiokaze_amd.devmem import DeviceArray, Stream, device_synchronize

    residual = 1e-6
    log(f"[bench] synthesising the batched workload: {loadgen.C4_GRIDS} grids")
    scenes = loadgen.c4_scenes(workers=min(16, os.cpu_count() or 1))
    cells = sum(sc.w * sc.h * sc.d for sc in scenes)
    log(f"[bench] sides {sorted(sc.w for sc in scenes)}; {cells / 1e6:.1f} M cells")
    dev = []
    for sc in scenes:
        dev.append(dict(fl=DeviceArray(sc.fluid), src=[DeviceArray(a) for a in (sc.u, sc.v, sc.wv)],
                        vel=[DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]))
    st = Stream()
    stream = st.handle
    device_synchronize()
    grids = [S.Context.device_grid(sc.w, sc.h, sc.d, sc.dt, b["fl"], S.F64, None, S.F32, *b["vel"])
             for sc, b in zip(scenes, dev)]
    ctx = S.Context(0)
    tune = 0 if args.no_autotune else 1
    prm_mixed = S.default_params(residual=residual, autotune=tune)
    prm_fp64 = S.default_params(residual=residual, precision=S.PREC_FP64, autotune=tune)

    def restore():
        for b in dev:
            for a, c in zip(b["vel"], b["src"]):
                a.copy_from(c, stream)

    def batched(prm):
        restore()
        return ctx.project_batch_device(grids, prm, stream=stream)

    def single(prm):
        out = []
        for g, b in zip(grids, dev):
            for a, c in zip(b["vel"], b["src"]):
                a.copy_from(c, stream)
            out.append(ctx.project_batch_device([g], prm, stream=stream)[0])
        return out

    def timed(fn, prm, n, label):
        for _ in range(max(1, args.warmup if fn is batched else 1)):
            fn(prm)
        st.synchronize()
        lat, res = [], None
        t0 = time.perf_counter()
        for _ in range(n):
            ts = time.perf_counter()
            res = fn(prm)
            lat.append((time.perf_counter() - ts) * 1e3)
        st.synchronize()
        el = time.perf_counter() - t0
        worst = max(r.reresid_inf for r in res)
        conv = all(r.converged for r in res)
        its = [r.iterations for r in res]
        log(f"[bench] {label}: {n} x {len(res)} solves in {el:.3f}s -> {n * len(res) / el:.1f} solves/s; "
            f"iterations {min(its)}..{max(its)}; worst reresid_inf {worst:.2e}; all converged {conv}")
        return n * len(res) / el, el / n * 1e3, lat, worst, conv

    v_b, ms_b, lat_b, worst_b, conv_b = timed(batched, prm_mixed, steps, "batched mixed")
    v_f, ms_f, _, worst_f, conv_f = timed(single, prm_fp64, single_rounds, "single FP64")
    stats, inst_steps = {}, 0
    if kernel_timing:
        ctx.set_kernel_timing(True)
        ctx.reset_kernel_stats()
        inst_steps = max(1, min(steps, 3))
        for _ in range(inst_steps):
            batched(prm_mixed)
        st.synchronize()
        stats = ctx.kernel_stats()
        ctx.set_kernel_timing(False)
    ctx.close()
    block = {
        "workload": (f"{loadgen.C4_GRIDS} sphere-drop cubes, sides from {list(loadgen.C4_SIDES)} (seed 3), "
                     f"{cells / 1e6:.1f} M cells; mgpcg mixed to {residual:g} (inf norm), one batched call per step"),
        "solves_per_s": round(v_b, 3), "ms_per_batch": round(ms_b, 3), "steps": steps,
        "worst_reresid_inf": worst_b, "all_converged": conv_b, "cells": int(cells),
        "sides": sorted(int(sc.w) for sc in scenes), "cell_updates_per_s": round(v_b / loadgen.C4_GRIDS * cells, 1),
        "single_fp64_solves_per_s": round(v_f, 3), "single_fp64_worst_reresid_inf": worst_f,
        "single_fp64_all_converged": conv_f, "speedup_vs_single_fp64": round(v_b / v_f, 3),
        "batch_latency_ms": {"p50": round(float(np.percentile(lat_b, 50)), 3),
                             "p99": round(float(np.percentile(lat_b, 99)), 3)},
    }
    return block, stats, inst_steps


def bench_batch(args):
    """--config c4: the batched workload as its own line."""
    block, stats, inst_steps = batched_block(args, args.steps, max(1, args.steps // 2),
                                             kernel_timing=not args.no_kernel_timing)
    out = {
        "metric": "batched pressure solves/sec (64 uneven grids to 1e-6 rel-res)",
        "value": block["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": block["ms_per_batch"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic", "config": block,
        "roofline": kernel_roofline(stats, inst_steps, "c4"), "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def served_block(args, n_requests):
    """The served workload: the Poisson-arrival stream (200 requests/s, seed 4) of 20-step
    time-stepped requests against the server (continuous batching, HBM-budget admission, warm start
    from resident pressure).  Returns the block dict."""
    import shiokaze_amd as S
    from shiokaze_amd import loadgen

    specs = loadgen.make_requests(n_requests, rate=200.0, seed=4)
    steps = args.c5_steps
    # library default controls: tuning on the first batch of each new shape class would land in
    # the served latencies
    prm = S.default_params(residual=1e-6, warm_start=1)
    cells = [s.w * s.h * s.d for s in specs]
    log(f"[bench] served: {len(specs)} requests x {steps} steps, mean {np.mean(cells) / 1e6:.2f} M cells, "
        f"arrivals over {specs[-1].t_arrival:.2f}s")
    # warm-up: a few requests through a separate server (code paths, allocations, graphs)
    with S.Server(0, executors=args.c5_executors) as srv:
        loadgen.run_served(srv, loadgen.make_requests(4, rate=1e6, seed=99), steps=2, params=prm)
    with S.Server(0, executors=args.c5_executors, max_batch_grids=64, hbm_budget_bytes=200 * 10 ** 9) as srv:
        res = loadgen.run_served(srv, specs, steps=steps, params=prm, log=log)
        st = srv.stats()
    ok = sum(1 for x in res["statuses"] if x == 0)
    value = res["solves"] / res["elapsed_s"]
    rl, sl = res["request_latency_ms"], res["solve_latency_ms"]
    offered = 200.0 * steps
    log(f"[bench] served: {res['solves']} solves ({ok} ok) in {res['elapsed_s']:.2f}s -> {value:.1f} solves/s; "
        f"request p50 {np.percentile(rl, 50):.0f} ms p99 {np.percentile(rl, 99):.0f} ms; "
        f"solve p50 {np.percentile(sl, 50):.1f} ms p99 {np.percentile(sl, 99):.1f} ms; server {st}")
    return {
        "workload": (f"{len(specs)} requests (Poisson arrivals, 200/s, seed 4) x {steps} steps, sides U[48,192] "
                     f"aspect <= 2, translated sphere-drop, warm start, mgpcg mixed to 1e-6"),
        "requests": len(specs), "requests_full_stream": 10000,
        "reduced": (f"the first {len(specs)} requests of the 10,000-request stream; the stream offers "
                    f"{offered:.0f} solves/s, above this device's capacity, so request latency is queueing "
                    f"and grows with the stream's length; solve latency (start of a request's batch to its "
                    f"completion) does not") if len(specs) < 10000 else "full stream",
        "steps_per_request": steps, "executors": args.c5_executors, "hbm_budget_bytes": 200 * 10 ** 9,
        "solves": res["solves"], "ok_solves": ok, "elapsed_s": round(res["elapsed_s"], 3),
        "solves_per_s": round(value, 3), "offered_solves_per_s": offered,
        "mean_cells": float(np.mean(cells)), "cell_updates_per_s": round(sum(cells) * steps / res["elapsed_s"], 1),
        "median_iterations": float(np.median(res["iterations"])) if res["iterations"] else None,
        "request_latency_ms": {"p50": round(float(np.percentile(rl, 50)), 3),
                               "p99": round(float(np.percentile(rl, 99)), 3)},
        "solve_latency_ms": {"p50": round(float(np.percentile(sl, 50)), 3),
                             "p99": round(float(np.percentile(sl, 99)), 3)},
        "server": st,
    }


def bench_serve(args):
    """--config c5: the served workload as its own line."""
    block = served_block(args, args.c5_requests)
    out = {
        "metric": "served pressure solves/sec (Poisson 200 req/s, 20-step requests, 1e-6 rel-res)",
        "value": block["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": block["solves"], "warmup": 8,
        "ms_per_step": round(block["elapsed_s"] / max(block["solves"], 1) * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic", "config": block,
        "latency_ms": {"p50": block["request_latency_ms"]["p50"], "p99": block["request_latency_ms"]["p99"],
                       "what": "request (arrival to last step done)",
                       "solve_p50": block["solve_latency_ms"]["p50"], "solve_p99": block["solve_latency_ms"]["p99"]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def accuracy_block(ctx, sc, prm, residual):
    """The returned pressure against the FP64 oracle: the reference's PCG with its MIC(0) factor
    applied, solved to 1e-10 on the oracle's own assembly of the same scene; the reference PCG's
    own answer at the benchmark tolerance (its iterate at the first iteration meeting it) for
    scale; the maximum post-projection divergence over the unknowns for the GPU result and for the
    oracle's update with the oracle pressure."""
    from oracle import oracle as O
    work = sc.copy()
    r, P, P64 = ctx.project(work, prm, pressure64=True)
    t0 = time.perf_counter()
    A = O.assemble(sc)
    x = O.gather(A, P64)
    bmax = float(np.abs(A["b"]).max())
    rel_res = float(np.abs(A["b"] - O.spmv(A, x)).max() / bmax)
    t1 = time.perf_counter()
    x_ref, it, rr, ok, x_snap, it_snap = O.pcg_snap(A, residual=1e-10, snap_residual=residual,
                                                    mic_mode=O.MIC_APPLIED)
    t_pcg = time.perf_counter() - t1
    nref = float(np.linalg.norm(x_ref))
    err = float(np.linalg.norm(x - x_ref) / nref)
    err_ref = float(np.linalg.norm(x_snap - x_ref) / nref)
    div_gpu, _ = O.divergence(work, A["index_map"], A["area"])
    ref = sc.copy()
    O.velocity_update(ref, O.scatter(A, x_ref, (sc.w, sc.h, sc.d)).astype(np.float32), A["area"], A["rho"])
    div_ref, _ = O.divergence(ref, A["index_map"], A["area"])
    log(f"[bench] accuracy: GPU residual {rel_res:.2e} on the oracle's assembly; pressure rel-L2 {err:.3e} "
        f"(reference PCG at {residual:g}: {err_ref:.3e}); max|div| GPU {div_gpu:.4e} oracle {div_ref:.4e} "
        f"(oracle MIC(0)-PCG {it} iterations, {t_pcg:.1f}s; check {time.perf_counter() - t0:.1f}s)")
    return {
        "oracle": f"FP64 MIC(0)-PCG (pcg_solver.h:89-295, factor applied) to 1e-10 relative residual: {it} iterations",
        "gpu_reresid_inf_on_oracle_assembly": rel_res, "pressure_rel_l2": err,
        "reference_pcg_rel_l2_at_tolerance": err_ref, "reference_pcg_iterations_to_tolerance": int(it_snap),
        "max_div_gpu": float(div_gpu), "max_div_oracle": float(div_ref),
        "div_delta_abs": float(abs(div_gpu - div_ref)), "div_delta_rel_b_inf": float(abs(div_gpu - div_ref) / bmax),
        "b_inf": bmax, "div_gate": "|max div_gpu - max div_oracle| <= 1e-5 ||b||_inf",
        "div_gate_met": bool(abs(div_gpu - div_ref) <= 1e-5 * bmax),
    }, int(it_snap)


def bench_post(args):
    """The rest of a time step around the projection, on the c3 scene's projected velocity and
    post-projection activity, all resident in HBM: velocity extrapolation over `--post-width`
    passes and the solid constraint (shkz_extrapolate_velocity_device), then MacCormack advection
    of the liquid level set and of the velocity (shkz_advect_scalar_device,
    shkz_advect_velocity_device) with dt at CFL 2, trilinear or (--weno) WENO sampling.  Each step
    restores the inputs from device copies (inside the timed region)."""
    import shiokaze_amd as S
    from shiokaze_amd.devmem import DeviceArray, Stream, device_synchronize
    kind_name, n, seed, residual = CONFIGS["c3"]
    width = args.post_width
    weno = bool(args.weno)
    log(f"[bench] post-projection step: c3 {n}^3 seed={seed}, width {width}, {'WENO' if weno else 'trilinear'}")
    sc = S.make_scene(S.SCENE_SPHERES_BOX, n, seed=seed)
    shapes = [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    d_fluid, d_solid = DeviceArray(sc.fluid), DeviceArray(sc.solid)
    vel = [DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]
    act = [DeviceArray(np.zeros(s, np.uint8)) for s in shapes]
    st = Stream()
    ctx = S.Context(0)
    r = ctx.project_device(n, n, n, sc.dt, d_fluid, S.F64, d_solid, S.F32, *vel,
                           params=S.default_params(residual=residual), stream=st.handle, active_out=act)
    st.synchronize()
    vel0_host = [v.download() for v in vel]
    vel0 = [DeviceArray(v) for v in vel0_host]
    act0_host = [a.download() for a in act]
    act0 = [DeviceArray(a) for a in act0_host]
    phi_host = sc.fluid.astype(np.float32)
    phi0 = DeviceArray(phi_host)
    phi = DeviceArray(phi_host)
    dt = 2.0 * sc.dx / max(float(np.abs(v).max()) for v in vel0_host)
    device_synchronize()

    def step():
        for a, b in zip(vel + act + [phi], vel0 + act0 + [phi0]):
            a.copy_from(b, st.handle)
        ctx.extrapolate_velocity_device(n, n, n, vel, act, solid=d_solid, solid_dtype=S.F32, width=width,
                                        stream=st.handle)
        ctx.advect_scalar_device(n, n, n, vel, act, phi, dt, stream=st.handle, weno=weno)
        ctx.advect_velocity_device(n, n, n, vel, act, phi, S.F32, dt, stream=st.handle, weno=weno)

    for _ in range(args.warmup):
        step()
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    st.synchronize()
    elapsed = time.perf_counter() - t0
    activated = int(sum(int(a.download().sum()) for a in act) - sum(int(a.sum()) for a in act0_host))
    stats = {}
    inst = 0
    if not args.no_kernel_timing:
        ctx.set_kernel_timing(True)
        ctx.reset_kernel_stats()
        inst = max(1, min(args.steps, 5))
        for _ in range(inst):
            step()
        st.synchronize()
        stats = ctx.kernel_stats()
        ctx.set_kernel_timing(False)
    roofline = kernel_roofline(stats, inst, "post")
    faces = sum(a * b * c for a, b, c in shapes)
    weno_fp64 = None
    if weno and stats:
        # FP64 roofline of the WENO kernels: one sample is 43 six-point interpolations (36 along x,
        # 6 along y, 1 along z) of 174 flops each (divisions counted once) plus the linear weights
        # of the three axes (36 flops); a pass samples each active entry whose velocity is not zero
        # and, backward, outside the narrow band -- counted here as every active face and every
        # cell, an upper bound
        wk = {k: v for k, v in stats.items() if k.endswith("_weno")}
        t_ms = sum(v["total_ms"] for v in wk.values()) / inst
        active_after = int(sum(int(a.download().sum()) for a in act))
        samples = 2 * (active_after + n * n * n)
        flops = samples * (43 * 174 + 36)
        weno_fp64 = {"bound": "fp64", "kernels": sorted(wk), "ms_per_step": round(t_ms, 4),
                     "flops_per_step_upper": int(flops), "achieved_upper": round(flops / (t_ms * 1e-3) / 1e12, 2),
                     "peak": 78.6, "unit": "TFLOP/s",
                     "frac_upper": round(flops / (t_ms * 1e-3) / 1e12 / 78.6, 4)}
    stage_ms = {}
    for k, v in stats.items():
        stage_ms[k] = round(v["total_ms"] / max(inst, 1), 4)
    interp = "WENO" if weno else "trilinear"
    out = {"metric": f"post-projection steps/sec @{n}^3 (extrapolation width {width}, solid constraint, "
                     f"level-set and velocity advection" + (", WENO)" if weno else ")"),
           "value": round(args.steps / elapsed, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{n}^3 ({kind_name}, seed {seed}) projected velocity: {width} extrapolation "
                                  f"passes, the solid constraint, MacCormack advection of the level set and the "
                                  f"velocity at CFL 2 ({interp} sampling); inputs restored on the device each step",
                      "faces": int(faces), "faces_activated": activated, "kernel_ms_per_step": stage_ms,
                      "active_faces_before": int(sum(int(a.sum()) for a in act0_host)),
                      "projection_iterations": int(r.iterations)},
           "roofline": roofline, "cpu_baseline": None}
    if weno_fp64:
        out["weno_fp64"] = weno_fp64
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        vh = [v.copy() for v in vel0_host]
        ah = [a.copy() for a in act0_host]
        ph = phi_host.copy()
        t = time.perf_counter()
        O.extrapolate(n, n, n, vh, ah, width)
        O.constrain_velocity(n, n, n, vh, ah, sc.solid)
        O.advect_scalar(n, n, n, vh, ah, ph, dt, weno=weno)
        O.advect_velocity(n, n, n, vh, ah, ph, dt, weno=weno)
        el = time.perf_counter() - t
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        out["cpu_baseline"] = {"value": round(1.0 / el, 4), "unit": "steps/s", "cores": threads, "kind": "port",
                               "sample": f"one whole step at {n}^3 by the oracle (C; the advection loops use "
                                         f"OpenMP), {el:.2f} s"}
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS) + ["c4", "c5", "post"])
    ap.add_argument("--post-width", type=int, default=3, help="post: extrapolation passes")
    ap.add_argument("--weno", action="store_true", help="post: WENO sampling in the advection")
    ap.add_argument("--c5-requests", type=int, default=1000, help="served: requests of the stream measured")
    ap.add_argument("--batched-steps", type=int, default=5, help="default line: batched calls timed")
    ap.add_argument("--no-extras", action="store_true",
                    help="default line: skip the FP64 rate, accuracy, batched and served blocks")
    ap.add_argument("--c5-steps", type=int, default=20, help="c5: steps per request")
    ap.add_argument("--c5-executors", type=int, default=2, help="c5: server executors")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--solver", default="mgpcg", choices=["mgpcg", "cg"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="library default multigrid controls")
    ap.add_argument("--mg", default="", help="explicit multigrid controls pre,post,coarse_pre,coarse_post,"
                                             "coarse_solve_sweeps,omega,inner_residual (implies --no-autotune)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy_mb:
        from shiokaze_amd.devmem import DeviceArray, Stream
        nb = args.calibrate_copy_mb << 20
        a = DeviceArray(shape=(nb,), dtype=np.uint8)
        b = DeviceArray(shape=(nb,), dtype=np.uint8)
        a.zero()
        st = Stream()
        for _ in range(5):
            b.copy_from(a, st.handle)
        st.synchronize()
        print(json.dumps({"calibration_copy_bytes": nb, "copies": 5}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the projection of one scene is a "
                                   "single-GPU kernel pipeline and no multi-GPU path is built; run with --gpus 1"}))
        sys.exit(2)

    if args.config == "c4":
        bench_batch(args)
        return
    if args.config == "c5":
        bench_serve(args)
        return
    if args.config == "post":
        bench_post(args)
        return

    import shiokaze_amd as S
    from shiokaze_amd.devmem import DeviceArray, Stream, device_synchronize

    kind_name, n, seed, residual = CONFIGS[args.config]
    kind = {"drop": S.SCENE_DROP, "drop_solids": S.SCENE_DROP_SOLIDS, "spheres_box": S.SCENE_SPHERES_BOX}[kind_name]
    log(f"[bench] synthesising {args.config} {n}^3 seed={seed}")
    sc = S.make_scene(kind, n, seed=seed)
    # inputs resident in HBM (same HIP runtime as the library)
    d_fluid = DeviceArray(sc.fluid)
    d_solid = DeviceArray(sc.solid) if sc.solid is not None else None
    src = [DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]
    vel = [DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]
    st = Stream()
    stream = st.handle
    device_synchronize()

    prec = S.PREC_MIXED if args.precision == "mixed" else S.PREC_FP64
    solver = S.SOLVER_MGPCG if args.solver == "mgpcg" else S.SOLVER_CG
    prm = S.default_params(residual=residual, precision=prec, solver=solver,
                           autotune=0 if (args.no_autotune or args.mg) else 1, **explicit_mg(args.mg))
    ctx = S.Context(0)

    def step():
        for a, b in zip(vel, src):
            a.copy_from(b, stream)
        return ctx.project_device(sc.w, sc.h, sc.d, sc.dt, d_fluid, S.F64, d_solid, S.F32, vel[0], vel[1], vel[2],
                                  params=prm, stream=stream)

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: it={r.iterations} outer={r.outer_iterations} reresid_inf={r.reresid_inf:.2e} "
            f"converged={r.converged} {r.t_projection:.2f} ms")
    st.synchronize()
    # ---- timed region: whole projections, graphs on, no per-kernel events
    lat = []
    results = []
    stages = []
    st.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        r = step()
        lat.append((time.perf_counter() - ts) * 1e3)
        results.append((r.iterations, r.outer_iterations, r.reresid_inf, r.converged))
        stages.append((r.t_solid_fluid_fractions, r.t_build_highres_linsystem, r.t_linsolve, r.t_update_velocity))
    st.synchronize()
    elapsed = time.perf_counter() - t0
    its = [x[0] for x in results]
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s; iterations {min(its)}..{max(its)}; "
        f"reresid_inf max {max(x[2] for x in results):.2e}; all converged {all(x[3] for x in results)}")
    # ---- instrumented pass over the same steps: every kernel bracketed by HIP events on the
    # library's stream (direct launches), for per-kernel durations and the roofline
    stats = {}
    inst_steps = 0
    if not args.no_kernel_timing:
        ctx.set_kernel_timing(True)
        ctx.reset_kernel_stats()
        inst_steps = max(1, min(args.steps, 5))
        t1 = time.perf_counter()
        for i in range(inst_steps):
            step()
        st.synchronize()
        log(f"[bench] instrumented pass: {inst_steps} steps, {(time.perf_counter() - t1) / inst_steps * 1e3:.2f} ms/step")
        stats = ctx.kernel_stats()
        ctx.set_kernel_timing(False)
    roofline = kernel_roofline(stats, inst_steps, args.config)

    value = args.steps / elapsed
    if roofline and stats:
        # the whole projection: every kernel's algorithmic bytes per step over the timed step
        step_bytes = sum(v["bytes"] for v in stats.values()) / inst_steps
        gbs = step_bytes / (elapsed / args.steps) / 1e9
        roofline["projection"] = {"algorithmic_bytes_per_step": int(step_bytes), "achieved": round(gbs, 1),
                                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                                  "what": "all kernels' algorithmic bytes per projection / wall time per projection"}
    out = {
        "metric": "pressure solves/sec @256^3 to 1e-6 rel-res" if args.config == "c3" else
                  f"pressure solves/sec @{n}^3 to {residual:g} rel-res",
        "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if prec == S.PREC_MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 liquid pressure projection ({kind_name}, seed {seed}), "
                               f"{args.solver} {args.precision} to {residual:g} relative residual (inf norm)",
                   "grid": [n, n, n], "unknowns": int(r.n_unknowns), "bricks": int(r.n_bricks),
                   "levels": int(r.n_levels), "iterations": int(np.median(its)),
                   "outer_iterations": int(r.outer_iterations), "reresid_inf": float(max(x[2] for x in results)),
                   "mg_controls": mg_controls(r),
                   "stage_ms_median": dict(zip(("fractions_labelling", "build", "solve", "update"),
                                               (round(float(x), 3) for x in np.median(np.array(stages), axis=0)))),
                   "cell_updates_per_s": round(value * n ** 3, 1)},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    mic_iterations = None
    if args.config == "c3" and not args.no_extras and prec == S.PREC_MIXED and solver == S.SOLVER_MGPCG:
        # the plain single-solve FP64 GPU path on the same scene (the comparison point of the
        # batched/served claim), same timing protocol
        prm64 = S.default_params(residual=residual, precision=S.PREC_FP64,
                                 autotune=0 if (args.no_autotune or args.mg) else 1, **explicit_mg(args.mg))
        saved = prm
        prm = prm64
        for _ in range(2):
            r64 = step()
        st.synchronize()
        t0 = time.perf_counter()
        n64 = 5
        for _ in range(n64):
            r64 = step()
        st.synchronize()
        el64 = time.perf_counter() - t0
        prm = saved
        out["fp64_single"] = {"solves_per_s": round(n64 / el64, 3), "ms_per_solve": round(el64 / n64 * 1e3, 3),
                              "iterations": int(r64.iterations), "reresid_inf": r64.reresid_inf,
                              "converged": bool(r64.converged), "mg_controls": mg_controls(r64),
                              "what": "plain FP64 MGPCG, one projection per call, same scene and timing"}
        log(f"[bench] FP64 single solve: {n64 / el64:.1f} solves/s ({r64.iterations} iterations)")
        log("[bench] accuracy against the FP64 oracle ...")
        out["accuracy"], mic_iterations = accuracy_block(ctx, sc, prm, residual)
    if not args.no_cpu_baseline:
        # iteration count of the reference's own (unpreconditioned) CG on this scene, from the GPU
        prm_cg = S.default_params(residual=residual, precision=S.PREC_FP64, solver=S.SOLVER_CG)
        for a, b in zip(vel, src):
            a.copy_from(b, stream)
        rcg = ctx.project_device(sc.w, sc.h, sc.d, sc.dt, d_fluid, S.F64, d_solid, S.F32, vel[0], vel[1], vel[2],
                                 params=prm_cg, stream=stream)
        log(f"[bench] reference CG iteration on GPU: {rcg.iterations} iterations, {rcg.t_projection:.1f} ms")
        out["config"]["reference_cg_iterations"] = int(rcg.iterations)
        out["config"]["gpu_reference_cg_ms"] = round(rcg.t_projection, 2)
        log("[bench] CPU baseline sample ...")
        out["cpu_baseline"] = cpu_baseline(sc, rcg.iterations, residual, mic_iterations)
    ctx.close()
    for a in [d_fluid, d_solid] + src + vel:
        if a is not None:
            a.free()
    if args.config == "c3" and not args.no_extras:
        out["batched"], _, _ = batched_block(args, args.batched_steps)
        if out.get("fp64_single"):
            out["batched"]["speedup_vs_c3_fp64_cells_per_s"] = round(
                out["batched"]["cell_updates_per_s"] / (out["fp64_single"]["solves_per_s"] * n ** 3), 3)
        out["served"] = served_block(args, args.c5_requests)
        out["latency_ms"]["served_request_p50"] = out["served"]["request_latency_ms"]["p50"]
        out["latency_ms"]["served_request_p99"] = out["served"]["request_latency_ms"]["p99"]
        out["latency_ms"]["served_solve_p50"] = out["served"]["solve_latency_ms"]["p50"]
        out["latency_ms"]["served_solve_p99"] = out["served"]["solve_latency_ms"]["p99"]
        out["latency_ms"]["what"] = ("p50/p99: closed loop of single projections of this scene; served_*: the "
                                     "served stream (request = arrival to last step, solve = one projection)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
