"""Batched projections: every grid of a batch gets exactly the result it gets alone.

A batch concatenates the grids' bricks into one tile table per multigrid level, but reductions,
multigrid depth, defect-correction state and the convergence test stay per grid, so the bar here
is bit-identity with the solo projection (pressure, velocities, iteration counts, residuals).
The solo projections themselves are checked against the oracle in test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def no_liquid(n):
    fluid = np.full((n, n, n), 0.5)
    return S.Scene(n, n, n, 1 / 120, fluid, None, *(np.ones(s, np.float32) for s in O.face_shapes(n, n, n)))


def mixed_batch():
    return [
        S.make_scene(S.SCENE_DROP, 24, seed=0),
        S.make_scene(S.SCENE_DROP_SOLIDS, 40, seed=1),
        S.make_scene(S.SCENE_RANDOM_DROP, 17, 30, 22, seed=3),
        S.make_scene(S.SCENE_DROP, 8, seed=4),            # one brick: depth 1
        no_liquid(16),                                     # no unknowns
        S.make_scene(S.SCENE_SPHERES_BOX, 48, seed=2),
        S.make_scene(S.SCENE_RANDOM_DROP, 5, 7, 3, seed=5),
    ]


def same(a, b):
    return a is None and b is None or np.array_equal(a, b)


def check_identical(rb, sb, P64b, solo):
    for g, (r, sc, P64) in enumerate(zip(rb, sb, P64b)):
        r1, sc1, P1 = solo[g]
        ctx_msg = f"grid {g} ({sc.w}x{sc.h}x{sc.d})"
        assert r.status == 0 and r1.status == 0, ctx_msg
        assert r.n_unknowns == r1.n_unknowns, ctx_msg
        assert r.n_levels == r1.n_levels, ctx_msg
        assert r.iterations == r1.iterations, (ctx_msg, r.iterations, r1.iterations)
        assert r.outer_iterations == r1.outer_iterations, ctx_msg
        assert r.converged == r1.converged, ctx_msg
        assert r.reresid == r1.reresid, (ctx_msg, r.reresid, r1.reresid)
        assert np.array_equal(P64, P1), ctx_msg
        for a, b in ((sc.u, sc1.u), (sc.v, sc1.v), (sc.wv, sc1.wv)):
            assert np.array_equal(a, b), ctx_msg


@pytest.mark.parametrize("mode", ["mixed", "fp64", "cg"])
def test_batch_equals_solo(ctx, mode):
    kw = {"mixed": dict(residual=1e-6), "fp64": dict(residual=1e-8, precision=S.PREC_FP64),
          "cg": dict(residual=1e-6, solver=S.SOLVER_CG, precision=S.PREC_FP64)}[mode]
    prm = S.default_params(**kw)
    scenes = mixed_batch()
    say(f"batch of {len(scenes)} ({mode}) vs solo")
    solo = []
    for sc in scenes:
        w = sc.copy()
        r, _, P64 = ctx.project(w, prm, pressure64=True)
        solo.append((r, w, P64))
    work = [sc.copy() for sc in scenes]
    rb, _, P64b = ctx.project_batch(work, prm, pressure64=True)
    check_identical(rb, work, P64b, solo)
    for r, (r1, _, _) in zip(rb, solo):
        assert r.converged


def test_batch_order_does_not_matter(ctx):
    say("batch order")
    prm = S.default_params(residual=1e-6)
    scenes = mixed_batch()
    w1 = [sc.copy() for sc in scenes]
    r1, _, P1 = ctx.project_batch(w1, prm, pressure64=True)
    order = [5, 2, 6, 0, 4, 1, 3]
    w2 = [scenes[i].copy() for i in order]
    r2, _, P2 = ctx.project_batch(w2, prm, pressure64=True)
    for k, i in enumerate(order):
        assert r2[k].iterations == r1[i].iterations
        assert np.array_equal(P2[k], P1[i])
        assert np.array_equal(w2[k].u, w1[i].u)


def test_batch_device_entry_matches_host_entry(ctx):
    say("batch device entry")
    from shiokaze_amd.devmem import DeviceArray, Stream
    prm = S.default_params(residual=1e-6)
    scenes = [S.make_scene(S.SCENE_DROP_SOLIDS, n, seed=n) for n in (16, 32, 24)]
    host = [sc.copy() for sc in scenes]
    rh, _, Ph = ctx.project_batch(host, prm, pressure64=True)
    bufs, grids = [], []
    for sc in scenes:
        b = dict(fl=DeviceArray(sc.fluid), so=DeviceArray(sc.solid), u=DeviceArray(sc.u), v=DeviceArray(sc.v),
                 w=DeviceArray(sc.wv), p=DeviceArray(shape=(sc.w, sc.h, sc.d), dtype=np.float64))
        bufs.append(b)
        grids.append(S.Context.device_grid(sc.w, sc.h, sc.d, sc.dt, b["fl"], S.F64, b["so"], S.F32, b["u"], b["v"],
                                           b["w"], pressure64=b["p"]))
    st = Stream()
    rd = ctx.project_batch_device(grids, prm, stream=st.handle)
    st.synchronize()
    for g, b in enumerate(bufs):
        assert rd[g].iterations == rh[g].iterations
        assert np.array_equal(b["p"].download(), Ph[g])
        assert np.array_equal(b["u"].download(), host[g].u)
        assert np.array_equal(b["w"].download(), host[g].wv)
        for a in b.values():
            a.free()
    st.destroy()


def test_batch_warm_start_scenes(ctx):
    """Two time steps of three scenes with warm start: batched in one context, one by one in
    another; the second step (warm-started from each scene's own resident pressure) agrees."""
    say("batch warm start")
    prm = S.default_params(residual=1e-6, warm_start=1)
    scenes = [S.make_scene(S.SCENE_RANDOM_DROP, n, seed=10 + n) for n in (20, 36, 28)]
    ids = [101, 102, 103]
    cb, cs = S.Context(0), S.Context(0)
    try:
        wb = [sc.copy() for sc in scenes]
        ws = [sc.copy() for sc in scenes]
        for step in range(2):
            rb, _, Pb = cb.project_batch(wb, prm, pressure64=True, scene_ids=ids)
            solo = [cs.project(ws[i], prm, pressure64=True, scene_id=ids[i]) for i in range(3)]
            for i in range(3):
                assert rb[i].iterations == solo[i][0].iterations, (step, i)
                assert np.array_equal(Pb[i], solo[i][2]), (step, i)
                assert np.array_equal(wb[i].v, ws[i].v), (step, i)
    finally:
        cb.close()
        cs.close()


def test_batch_rejects_duplicate_scene_ids(ctx):
    say("duplicate scene ids")
    scenes = [S.make_scene(S.SCENE_DROP, 16, seed=1), S.make_scene(S.SCENE_DROP, 16, seed=2)]
    rb, _, _ = ctx.project_batch(scenes, S.default_params(), scene_ids=[7, 7], check=False)
    assert rb[0].status == S.ERR_INVALID and rb[1].status == S.ERR_INVALID
    assert "twice" in S.last_error()
