"""The projection at the benchmark configurations' full sizes (128^3 with solids, 256^3), through
properties that do not depend on size:

  * the FP64 residual of the returned pressure, measured on the oracle's own assembly of the
    system (CSR, FP64), meets the requested tolerance (infinity norm relative to b);
  * the returned float pressure is the FP64 pressure rounded;
  * the updated velocities are bit-identical to the oracle's velocity update applied to the GPU's
    pressure (every face of the box);
  * a second projection of the same input gives bit-identical results (determinism at scale);
  * at 256^3, the post-projection step (extrapolation of width 3 with the solid constraint, then
    MacCormack advection of the level set and of the velocity at CFL 2) is bit-identical to the
    oracle's on the same projected velocity.

Accuracy against the FP64 oracle, the reference's PCG (pcg_solver.h:246-295 with its MIC(0) factor
applied, pcg_solver.h:89-223) solved to 1e-10 relative residual on the same system
(test_accuracy_against_fp64_oracle):

  * C2 (128^3 with solids) in plain FP64 MGPCG to 1e-8, and C3 (256^3) in mixed precision to 1e-6;
  * the pressure's relative-L2 error against the 1e-10 oracle solution must not exceed ten times
    the error of the reference PCG's own answer at the same tolerance (the oracle's iterate at the
    first iteration that meets it), so the GPU's answer is as good as the reference's;
  * the maximum post-projection divergence over the unknowns (div = sum of sgn*area*u/dx over a
    cell's faces, the negated assembly rhs) is reported for the GPU result and for the oracle's
    update with the oracle pressure, absolutely and relative to ||b||_inf (the pre-projection
    divergence), and the difference must stay within 1e-5 of ||b||_inf: after the update the
    divergence equals minus the residual, so a relative residual of 1e-6 bounds it by 1e-6 ||b||_inf
    up to the float storage of the velocities.
"""
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FULL = {
    # name: (scene kind, n, seed, residual) as the benchmark configurations set them
    "c2_128": (S.SCENE_DROP_SOLIDS, 128, 1, 1e-8),
    "c3_256": (S.SCENE_SPHERES_BOX, 256, 2, 1e-6),
}


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", list(FULL))
def test_full_size_projection(ctx, name):
    kind, n, seed, residual = FULL[name]
    say(f"full size {name}")
    sc = S.make_scene(kind, n, seed=seed)
    prm = S.default_params(residual=residual, precision=S.PREC_MIXED)
    work = sc.copy()
    r, P, P64 = ctx.project(work, prm, pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    assert np.array_equal(P, P64.astype(np.float32))

    A = O.assemble(sc)
    assert r.n_unknowns == A["n"]
    x = O.gather(A, P64)
    resid = A["b"] - O.spmv(A, x)
    rel = np.abs(resid).max() / np.abs(A["b"]).max()
    say(f"{name}: n={A['n']} iterations={r.iterations} residual={rel:.3e}")
    assert rel <= residual * 1.0001, rel

    ref = sc.copy()
    O.velocity_update(ref, P, A["area"], A["rho"])
    for a, b in ((work.u, ref.u), (work.v, ref.v), (work.wv, ref.wv)):
        assert np.array_equal(a, b)

    again = sc.copy()
    r2, P2, P64b = ctx.project(again, prm, pressure64=True)
    assert r2.iterations == r.iterations
    assert np.array_equal(P64b, P64)
    for a, b in ((again.u, work.u), (again.v, work.v), (again.wv, work.wv)):
        assert np.array_equal(a, b)


def test_full_size_post_step(ctx):
    say("full size post step 256")
    sc = S.make_scene(S.SCENE_SPHERES_BOX, 256, seed=2)
    work = sc.copy()
    shapes = O.face_shapes(sc.w, sc.h, sc.d)
    act = [np.zeros(s, np.uint8) for s in shapes]
    r, _, _ = ctx.project(work, S.default_params(residual=1e-6), active_out=act)
    assert r.status == 0 and r.converged
    vel = [work.u, work.v, work.wv]
    # the GPU and the oracle start from the same projected velocity and activity
    vg, ag = [v.copy() for v in vel], [a.copy() for a in act]
    vo, ao = [v.copy() for v in vel], [a.copy() for a in act]
    ctx.extrapolate_velocity(sc.w, sc.h, sc.d, vg, ag, solid=sc.solid, width=3)
    O.extrapolate(sc.w, sc.h, sc.d, vo, ao, 3)
    O.constrain_velocity(sc.w, sc.h, sc.d, vo, ao, sc.solid)
    for a in range(3):
        assert np.array_equal(ag[a], ao[a]), a
        assert np.array_equal(vg[a].view(np.uint32), vo[a].view(np.uint32)), a
    dt = 2.0 * sc.dx / max(float(np.abs(v).max()) for v in vg)  # CFL 2
    phi_g = sc.fluid.astype(np.float32)
    phi_o = phi_g.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, vg, ag, phi_g, dt)
    O.advect_scalar(sc.w, sc.h, sc.d, vo, ao, phi_o, dt)
    assert np.array_equal(phi_g.view(np.uint32), phi_o.view(np.uint32))
    ctx.advect_velocity(sc.w, sc.h, sc.d, vg, ag, phi_g, dt)
    O.advect_velocity(sc.w, sc.h, sc.d, vo, ao, phi_o, dt)
    for a in range(3):
        assert np.array_equal(vg[a].view(np.uint32), vo[a].view(np.uint32)), a


ACCURACY = {
    # name: (scene kind, n, seed, residual, precision)
    "c2_128_fp64": (S.SCENE_DROP_SOLIDS, 128, 1, 1e-8, S.PREC_FP64),
    "c3_256_mixed": (S.SCENE_SPHERES_BOX, 256, 2, 1e-6, S.PREC_MIXED),
}


@pytest.mark.parametrize("name", list(ACCURACY))
def test_accuracy_against_fp64_oracle(ctx, name):
    kind, n, seed, residual, prec = ACCURACY[name]
    say(f"accuracy {name}")
    sc = S.make_scene(kind, n, seed=seed)
    work = sc.copy()
    r, P, P64 = ctx.project(work, S.default_params(residual=residual, precision=prec), pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    A = O.assemble(sc)
    assert r.n_unknowns == A["n"]
    x = O.gather(A, P64)
    bmax = np.abs(A["b"]).max()
    rel_res = np.abs(A["b"] - O.spmv(A, x)).max() / bmax
    assert rel_res <= residual * 1.0001, rel_res
    # the FP64 oracle to 1e-10, and the reference PCG's own answer at the GPU's tolerance
    x_ref, it, rr, ok, x_snap, it_snap = O.pcg_snap(A, residual=1e-10, snap_residual=residual,
                                                    mic_mode=O.MIC_APPLIED)
    assert ok and it_snap > 0
    nref = np.linalg.norm(x_ref)
    err = np.linalg.norm(x - x_ref) / nref
    err_ref = np.linalg.norm(x_snap - x_ref) / nref
    # post-projection divergence: GPU result, and the oracle's update with the oracle pressure
    div_gpu, _ = O.divergence(work, A["index_map"], A["area"])
    ref = sc.copy()
    O.velocity_update(ref, O.scatter(A, x_ref, (n, n, n)).astype(np.float32), A["area"], A["rho"])
    div_ref, _ = O.divergence(ref, A["index_map"], A["area"])
    ddiv = abs(div_gpu - div_ref)
    say(f"{name}: unknowns {A['n']}, GPU {r.iterations} iterations to {rel_res:.2e}; oracle MIC(0)-PCG {it} "
        f"iterations to {rr:.1e} ({it_snap} to {residual:g}); pressure rel-L2 {err:.3e} (reference PCG at "
        f"{residual:g}: {err_ref:.3e}); max|div| GPU {div_gpu:.4e} oracle {div_ref:.4e} |delta| {ddiv:.4e} "
        f"= {ddiv / bmax:.3e} of ||b||_inf {bmax:.4e}")
    assert err <= 10.0 * err_ref, (err, err_ref)
    assert ddiv <= 1e-5 * bmax, (ddiv, bmax)
