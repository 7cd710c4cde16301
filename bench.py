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
C4_SIDES = (32, 48, 64, 96, 128, 160)
C4_GRIDS = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


SYMBOL = {"jacobi_L0": "k_jacobi4<1>", "jacobi_dot_L0": "k_jacobi4_dot",
          "residual_L0": "k_residual4<1>", "smooth2_L0": "k_smooth2<float, 1>"}


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


def cpu_baseline(scene, n_iter_needed, residual):
    """Reference-faithful CPU projection (the oracle restating macpressuresolver3 + pcg with MIC(0)
    computed and discarded), timed on a bounded sample: the full fractions/labelling/assembly and
    velocity update, plus 20 and 40 CG iterations to separate the per-iteration cost; the solve is
    extrapolated to the iteration count the same CG iteration needs on this scene."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    A = O.assemble(scene)
    t_build = time.perf_counter() - t0
    t1 = time.perf_counter()
    O.pcg(A, residual=residual, max_iterations=20, mic_mode=O.MIC_FAITHFUL)
    t20 = time.perf_counter() - t1
    t1 = time.perf_counter()
    O.pcg(A, residual=residual, max_iterations=40, mic_mode=O.MIC_FAITHFUL)
    t40 = time.perf_counter() - t1
    per_it = max((t40 - t20) / 20.0, 1e-9)
    fixed = max(t20 - 20 * per_it, 0.0)
    P = np.zeros((scene.w, scene.h, scene.d), np.float32)
    work = scene.copy()
    t2 = time.perf_counter()
    O.velocity_update(work, P, A["area"], A["rho"])
    t_upd = time.perf_counter() - t2
    t_solve = fixed + n_iter_needed * per_it
    total = t_build + t_solve + t_upd
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": (f"oracle (reference-faithful: MIC(0) factored and discarded, serial CG, threaded "
                   f"fractions/assembly/update on {threads} threads) on the same {scene.w}x{scene.h}x{scene.d} scene: "
                   f"build {t_build:.2f}s + update {t_upd:.2f}s measured, CG {per_it * 1e3:.1f} ms/iteration "
                   f"(from 20 and 40 iterations) x {n_iter_needed} iterations (count of the same CG iteration run "
                   f"to {residual:g} on the GPU) + factor {fixed:.2f}s = {total:.1f}s per projection"),
    }


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
    """Params fields of an explicit --mg pre,post,coarse_pre,coarse_post,coarse_sweeps,omega,eta."""
    if not spec:
        return {}
    v = spec.split(",")
    return dict(mg_pre_sweeps=int(v[0]), mg_post_sweeps=int(v[1]), mg_coarse_pre_sweeps=int(v[2]),
                mg_coarse_post_sweeps=int(v[3]), mg_coarse_sweeps=int(v[4]), mg_omega=float(v[5]),
                inner_residual=float(v[6]))


def mg_controls(r):
    """The multigrid controls a result reports (autotuned or default)."""
    return {"pre": r.mg_pre_sweeps, "post": r.mg_post_sweeps, "coarse_pre": r.mg_coarse_pre_sweeps,
            "coarse_post": r.mg_coarse_post_sweeps, "coarse_solve_sweeps": r.mg_coarse_sweeps,
            "omega": r.mg_omega, "inner_residual": r.inner_residual, "tuned": r.tuned}


def bench_batch(args):
    """C4: 64 uneven grids projected in one batched call per step (mixed precision, 1e-6), against
    the plain single-solve FP64 GPU path over the same grids one by one."""
    import shiokaze_amd as S
    from shiokaze_amd.devmem import DeviceArray, Stream, device_synchronize

    residual = 1e-6
    log(f"[bench] synthesising c4: {C4_GRIDS} grids")
    from shiokaze_amd import loadgen
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

    def timed(fn, prm, steps, label):
        for _ in range(args.warmup):
            fn(prm)
        st.synchronize()
        lat, res = [], None
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            res = fn(prm)
            lat.append((time.perf_counter() - ts) * 1e3)
        st.synchronize()
        el = time.perf_counter() - t0
        worst = max(r.reresid_inf for r in res)
        conv = all(r.converged for r in res)
        its = [r.iterations for r in res]
        log(f"[bench] {label}: {steps} x {len(res)} solves in {el:.3f}s -> {steps * len(res) / el:.1f} solves/s; "
            f"iterations {min(its)}..{max(its)}; worst reresid_inf {worst:.2e}; all converged {conv}")
        log(f"[bench] {label}: controls {mg_controls(res[-1])}")
        return steps * len(res) / el, el / steps * 1e3, lat, worst, conv

    v_b, ms_b, lat_b, worst_b, conv_b = timed(batched, prm_mixed, args.steps, "batched mixed")
    v_f, ms_f, _, worst_f, conv_f = timed(single, prm_fp64, max(1, args.steps // 2), "single FP64")
    v_m, _, _, _, _ = timed(single, prm_mixed, max(1, args.steps // 2), "single mixed")
    v_bf, _, _, _, _ = timed(batched, prm_fp64, max(1, args.steps // 2), "batched FP64")
    stats, inst_steps = {}, 0
    if not args.no_kernel_timing:
        ctx.set_kernel_timing(True)
        ctx.reset_kernel_stats()
        inst_steps = max(1, min(args.steps, 3))
        for _ in range(inst_steps):
            batched(prm_mixed)
        st.synchronize()
        stats = ctx.kernel_stats()
        ctx.set_kernel_timing(False)
    roofline = kernel_roofline(stats, inst_steps, "c4")
    ctx.close()
    out = {
        "metric": f"batched pressure solves/sec ({C4_GRIDS} uneven grids to 1e-6 rel-res)",
        "value": round(v_b, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_b, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{C4_GRIDS} sphere-drop cubes, sides from {list(C4_SIDES)} (seed 3), mgpcg mixed "
                               f"to {residual:g} relative residual (inf norm), one batched call per step",
                   "grids": C4_GRIDS, "cells": int(cells), "sides": sorted(int(sc.w) for sc in scenes),
                   "worst_reresid_inf": worst_b, "all_converged": conv_b,
                   "cell_updates_per_s": round(v_b / C4_GRIDS * cells, 1)},
        "single_fp64_solves_per_s": round(v_f, 3), "single_fp64_worst_reresid_inf": worst_f,
        "single_fp64_all_converged": conv_f, "single_mixed_solves_per_s": round(v_m, 3),
        "batched_fp64_solves_per_s": round(v_bf, 3), "speedup_vs_single_fp64": round(v_b / v_f, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat_b, 50)), 3),
                       "p99": round(float(np.percentile(lat_b, 99)), 3)},
        "roofline": roofline, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """C5: Poisson-arrival stream of time-stepped projection requests against the server
    (continuous batching, HBM-budget admission, warm start from resident pressure)."""
    import shiokaze_amd as S
    from shiokaze_amd import loadgen

    specs = loadgen.make_requests(args.c5_requests, rate=200.0, seed=4)
    steps = args.c5_steps
    # library default controls: tuning on the first batch of each new shape class would land in
    # the served latencies
    prm = S.default_params(residual=1e-6, warm_start=1)
    cells = [s.w * s.h * s.d for s in specs]
    log(f"[bench] c5: {len(specs)} requests x {steps} steps, mean {np.mean(cells) / 1e6:.2f} M cells, "
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
    log(f"[bench] c5: {res['solves']} solves ({ok} ok) in {res['elapsed_s']:.2f}s -> {value:.1f} solves/s; "
        f"request p50 {np.percentile(rl, 50):.0f} ms p99 {np.percentile(rl, 99):.0f} ms; "
        f"solve p50 {np.percentile(sl, 50):.1f} ms p99 {np.percentile(sl, 99):.1f} ms; server {st}")
    out = {
        "metric": "served pressure solves/sec (Poisson 200 req/s, 20-step requests, 1e-6 rel-res)",
        "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": res["solves"], "warmup": 8,
        "ms_per_step": round(res["elapsed_s"] / max(res["solves"], 1) * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(specs)} requests (Poisson arrivals, 200/s, seed 4) x {steps} steps, sides "
                               f"U[48,192] aspect <= 2, translated sphere-drop, warm start, mgpcg mixed to 1e-6",
                   "requests": len(specs), "steps_per_request": steps, "executors": args.c5_executors,
                   "hbm_budget_bytes": 200 * 10 ** 9, "mean_cells": float(np.mean(cells)),
                   "ok_solves": ok, "cell_updates_per_s": round(sum(cells) * steps / res["elapsed_s"], 1),
                   "median_iterations": float(np.median(res["iterations"])) if res["iterations"] else None,
                   "server": st},
        "latency_ms": {"p50": round(float(np.percentile(rl, 50)), 3), "p99": round(float(np.percentile(rl, 99)), 3),
                       "what": "request (arrival to last step done)",
                       "solve_p50": round(float(np.percentile(sl, 50)), 3),
                       "solve_p99": round(float(np.percentile(sl, 99)), 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_post(args):
    """The rest of a time step around the projection, on the c3 scene's projected velocity and
    post-projection activity, all resident in HBM: velocity extrapolation over `--post-width`
    passes and the solid constraint (shkz_extrapolate_velocity_device), then MacCormack advection
    of the liquid level set and of the velocity (shkz_advect_scalar_device,
    shkz_advect_velocity_device) with dt at CFL 2.  Each step restores the inputs from device
    copies (inside the timed region)."""
    import shiokaze_amd as S
    from shiokaze_amd.devmem import DeviceArray, Stream, device_synchronize
    kind_name, n, seed, residual = CONFIGS["c3"]
    width = args.post_width
    log(f"[bench] post-projection step: c3 {n}^3 seed={seed}, width {width}")
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
        ctx.advect_scalar_device(n, n, n, vel, act, phi, dt, stream=st.handle)
        ctx.advect_velocity_device(n, n, n, vel, act, phi, S.F32, dt, stream=st.handle)

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
    stage_ms = {}
    for k, v in stats.items():
        stage_ms[k] = round(v["total_ms"] / max(inst, 1), 4)
    out = {"metric": f"post-projection steps/sec @{n}^3 (extrapolation width {width}, solid constraint, "
                     f"level-set and velocity advection)",
           "value": round(args.steps / elapsed, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{n}^3 ({kind_name}, seed {seed}) projected velocity: {width} extrapolation "
                                  f"passes, the solid constraint, MacCormack advection of the level set and the "
                                  f"velocity at CFL 2; inputs restored on the device each step",
                      "faces": int(faces), "faces_activated": activated, "kernel_ms_per_step": stage_ms,
                      "active_faces_before": int(sum(int(a.sum()) for a in act0_host)),
                      "projection_iterations": int(r.iterations)},
           "roofline": roofline, "cpu_baseline": None}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        vh = [v.copy() for v in vel0_host]
        ah = [a.copy() for a in act0_host]
        ph = phi_host.copy()
        t = time.perf_counter()
        O.extrapolate(n, n, n, vh, ah, width)
        O.constrain_velocity(n, n, n, vh, ah, sc.solid)
        O.advect_scalar(n, n, n, vh, ah, ph, dt)
        O.advect_velocity(n, n, n, vh, ah, ph, dt)
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
    ap.add_argument("--c5-requests", type=int, default=200, help="c5: requests in the sample")
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
        out["cpu_baseline"] = cpu_baseline(sc, rcg.iterations, residual)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
