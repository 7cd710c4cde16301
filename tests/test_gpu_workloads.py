"""The batched and served workloads at their real shapes, checked grid by grid against the oracle.

Batched (64 cubes, sides drawn from {32,48,64,96,128,160} with seed 3, each a seeded sphere-drop
with its own N(0,1) velocities; one batched call, mixed precision to 1e-6):
  * every grid's FP64 residual on the oracle's own assembly meets 1e-6 (infinity norm relative to
    b, the reference's criterion, pcg_solver.h:252-283);
  * every grid's velocities are bit-identical to the oracle's velocity update
    (macpressuresolver3.cpp:252-268) applied to the GPU pressure;
  * every grid's pressure, velocities, iteration count and residual are bit-identical to projecting
    that grid alone.

Served (a prefix of the Poisson stream: rate 200/s, seed 4, sides U[48,192], aspect <= 2, 20 steps
per request with the liquid translated each step and warm start from the resident pressure,
macpressuresolver3.cpp:221-242), through a two-executor server with continuous batching:
  * every step of the kept requests equals the same step through shkz_project_device on one
    context, bit for bit;
  * at the first and last step of each kept request, the FP64 residual on the oracle's assembly of
    that step's inputs meets 1e-6 relative to the warm-started system's right-hand side
    b - A p_prev, the vector the reference's solver sees (:221-230).
"""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O
from shiokaze_amd import loadgen
from shiokaze_amd.devmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_batched_workload_per_grid_against_oracle():
    say("batched workload: 64 grids, seed 3")
    scenes = loadgen.c4_scenes(workers=8)
    assert len(scenes) == 64 and sorted({sc.w for sc in scenes}) == sorted(loadgen.C4_SIDES)
    prm = S.default_params(residual=1e-6)
    work = [sc.copy() for sc in scenes]
    with S.Context(0) as ctx:
        res, Ps, P64s = ctx.project_batch(work, prm, pressure64=True)
        assert all(r.status == 0 and r.converged for r in res)
        assert all(r.batch_size == 64 for r in res)
        worst = 0.0
        for g, (sc, r) in enumerate(zip(scenes, res)):
            A = O.assemble(sc)
            assert r.n_unknowns == A["n"], g
            if A["n"]:
                x = O.gather(A, P64s[g])
                rel = np.abs(A["b"] - O.spmv(A, x)).max() / np.abs(A["b"]).max()
                assert rel <= 1e-6 * 1.0001, (g, rel)
                worst = max(worst, rel)
            ref = sc.copy()
            O.velocity_update(ref, Ps[g], A["area"], A["rho"])
            for a, b in ((work[g].u, ref.u), (work[g].v, ref.v), (work[g].wv, ref.wv)):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), g
        say(f"  all 64 grids within 1e-6 on the oracle's assembly (worst {worst:.2e}); velocities bit-identical")
        for g, sc in enumerate(scenes):
            solo = sc.copy()
            r1, P1, P64_1 = ctx.project(solo, prm, pressure64=True)
            assert r1.iterations == res[g].iterations and r1.reresid_inf == res[g].reresid_inf, g
            assert np.array_equal(P64_1, P64s[g]), g
            for a, b in ((solo.u, work[g].u), (solo.v, work[g].v), (solo.wv, work[g].wv)):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), g
        say("  every grid equals its solo projection bit for bit")


def download_inputs(spec, step, stream):
    nbytes, off = loadgen.layout(spec)
    buf = DeviceArray(shape=(nbytes,), dtype=np.uint8)
    ptrs = {k: buf.data_ptr() + o for k, o in off.items()}
    loadgen.step_inputs(spec, step, ptrs, stream.handle)
    stream.synchronize()
    raw = buf.download()
    buf.free()
    w, h, d = spec.w, spec.h, spec.d

    def arr(key, shape):
        o = off[key]
        return raw[o:o + 4 * int(np.prod(shape))].view(np.float32).reshape(shape).copy()

    return SimpleNamespace(w=w, h=h, d=d, dt=1.0 / 120.0, fluid=arr("fluid", (w, h, d)), solid=None,
                           u=arr("u", (w + 1, h, d)), v=arr("v", (w, h + 1, d)), wv=arr("w", (w, h, d + 1)))


def test_served_stream_prefix_against_direct_and_oracle():
    steps = 20
    specs = loadgen.make_requests(12, rate=200.0, seed=4)
    # the stream's first requests (the generator draws sequentially, so these open every longer stream)
    assert [(s.w, s.h, s.d) for s in specs[:3]] == [(s.w, s.h, s.d) for s in loadgen.make_requests(3, seed=4)]
    keep = {0, 5, 11}
    prm = S.default_params(residual=1e-6, warm_start=1)
    say(f"served stream prefix: {len(specs)} requests x {steps} steps, sides "
        f"{[(s.w, s.h, s.d) for s in specs]}")
    with S.Server(0, executors=2, max_batch_grids=64, hbm_budget_bytes=200 * 10 ** 9) as srv:
        out = loadgen.run_served(srv, specs, steps=steps, params=prm, keep_pressure=keep)
        st = srv.stats()
    say(f"  {out['solves']} solves, server {st}")
    assert out["solves"] == len(specs) * steps and all(s == 0 for s in out["statuses"])
    assert st["batches"] < st["batched_requests"], "no two requests shared a batch"
    stream = Stream()
    ctx = S.Context(0)
    try:
        for idx in sorted(keep):
            spec = specs[idx]
            served = out["pressures"][idx]
            nbytes, off = loadgen.layout(spec)
            buf = DeviceArray(shape=(nbytes,), dtype=np.uint8)
            ptrs = {k: buf.data_ptr() + o for k, o in off.items()}
            p64 = DeviceArray(shape=(spec.w, spec.h, spec.d), dtype=np.float64)
            for t in range(steps):
                loadgen.step_inputs(spec, t, ptrs, stream.handle)
                r = ctx.project_device(spec.w, spec.h, spec.d, 1.0 / 120.0, ptrs["fluid"], S.F32, None, S.F32,
                                       ptrs["u"], ptrs["v"], ptrs["w"], params=prm, pressure64=p64,
                                       scene_id=1000 + idx, stream=stream.handle)
                stream.synchronize()
                assert r.status == 0 and r.converged, (idx, t)
                assert np.array_equal(p64.download(), served[t]), (idx, t)
            buf.free()
            p64.free()
            for t in (0, steps - 1):
                sc = download_inputs(spec, t, stream)
                A = O.assemble(sc)
                x = O.gather(A, served[t])
                b_eff = A["b"] if t == 0 else A["b"] - O.spmv(A, O.gather(A, served[t - 1]))
                rel = np.abs(A["b"] - O.spmv(A, x)).max() / np.abs(b_eff).max()
                say(f"  request {idx} {spec.w}x{spec.h}x{spec.d} step {t}: unknowns {A['n']}, residual {rel:.2e} "
                    f"of the warm-started rhs")
                assert rel <= 1e-6 * 1.0001, (idx, t, rel)
            ctx.release_scene(1000 + idx)
    finally:
        ctx.close()
        stream.destroy()
