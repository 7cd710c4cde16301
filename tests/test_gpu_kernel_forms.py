"""Alternative kernel forms of the same solver arithmetic give the same bits.

* Coarse levels smooth either with one launch per sweep (plus separate residual and prolongation
  launches) or with several sweeps in one launch in LDS; which form a level gets depends on its
  size, and so on the batch a grid is in.  Both forms must give bit-identical projections, so that
  a grid's result never depends on its batch (forced here with SHKZ_MULTI_MAX_BRICKS: 0 = one
  launch per sweep everywhere, a large limit = multi-sweep launches on every coarse level).
* The FP32 Krylov iteration runs either with the direction update fused into the operator apply
  and the first V-cycle sweep fused into the residual update, or with separate kernels
  (SHKZ_NO_FUSE=1); both must give bit-identical projections, alone and in a batch whose grids end
  on different levels, with several V-cycle shapes of the brick-local smoother.
"""
import os
import sys

import numpy as np
import pytest

import shiokaze_amd as S

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def project_with_env(sc, prm, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with S.Context(0) as ctx:
            work = sc.copy()
            r, P, P64 = ctx.project(work, prm, pressure64=True)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert r.status == 0 and r.converged, r.as_dict()
    return r, P64, work


def assert_same(a, b):
    ra, Pa, wa = a
    rb, Pb, wb = b
    assert ra.iterations == rb.iterations and ra.outer_iterations == rb.outer_iterations
    assert ra.reresid_inf == rb.reresid_inf
    assert np.array_equal(Pa.view(np.uint64), Pb.view(np.uint64))
    for x, y in ((wa.u, wb.u), (wa.v, wb.v), (wa.wv, wb.wv)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


CASES = {
    # name: (scene kind, n, seed, precision)
    "drop_solids_96_mixed": (S.SCENE_DROP_SOLIDS, 96, 1, S.PREC_MIXED),
    "spheres_box_128_mixed": (S.SCENE_SPHERES_BOX, 128, 2, S.PREC_MIXED),
    "drop_80_fp64": (S.SCENE_DROP, 80, 0, S.PREC_FP64),
}


@pytest.mark.parametrize("name", list(CASES))
def test_multi_sweep_and_per_sweep_smoothing_agree(name):
    kind, n, seed, prec = CASES[name]
    say(f"smoothing forms {name}")
    sc = S.make_scene(kind, n, seed=seed)
    prm = S.default_params(residual=1e-6, precision=prec)
    per_sweep = project_with_env(sc, prm, {"SHKZ_MULTI_MAX_BRICKS": "0"})
    multi = project_with_env(sc, prm, {"SHKZ_MULTI_MAX_BRICKS": "1000000"})
    default = project_with_env(sc, prm, {})
    say(f"  {name}: {per_sweep[0].iterations} iterations, reresid {per_sweep[0].reresid_inf:.3e}")
    assert_same(per_sweep, multi)
    assert_same(per_sweep, default)


@pytest.mark.parametrize("name", ["drop_solids_96_mixed", "spheres_box_128_mixed"])
def test_fused_and_separate_kernels_agree(name):
    kind, n, seed, prec = CASES[name]
    say(f"fused kernel forms {name}")
    sc = S.make_scene(kind, n, seed=seed)
    prm = S.default_params(residual=1e-6, precision=prec)
    fused = project_with_env(sc, prm, {})
    separate = project_with_env(sc, prm, {"SHKZ_NO_FUSE": "1"})
    assert_same(fused, separate)


@pytest.mark.parametrize("mg", [(2, 2, 2, 2), (1, 1, 1, 1), (2, 1, 1, 2)])
def test_fused_and_separate_kernels_agree_in_a_batch(mg):
    """Grids of 8^3 to 112^3 cells end on levels 0 to 4; the brick-local smoother with the
    given pre/post applications on the finest and the coarse levels."""
    say(f"fused kernel forms, batch, V-cycle {mg}")
    scenes = [S.make_scene(S.SCENE_DROP_SOLIDS, 112, seed=1), S.make_scene(S.SCENE_RANDOM_DROP, 17, 30, 22, seed=3),
              S.make_scene(S.SCENE_DROP, 8, seed=4), S.make_scene(S.SCENE_SPHERES_BOX, 64, seed=2),
              S.make_scene(S.SCENE_DROP, 40, seed=0)]
    prm = S.default_params(residual=1e-6, precision=S.PREC_MIXED)
    prm.mg_smoother = S.SMOOTH_BRICK
    prm.mg_pre_sweeps, prm.mg_post_sweeps, prm.mg_coarse_pre_sweeps, prm.mg_coarse_post_sweeps = mg
    out = []
    for env in ({}, {"SHKZ_NO_FUSE": "1"}):
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            with S.Context(0) as ctx:
                work = [sc.copy() for sc in scenes]
                res, _, P64 = ctx.project_batch(work, prm, pressure64=True)
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        out.append((res, P64, work))
    (ra, Pa, wa), (rb, Pb, wb) = out
    for g in range(len(scenes)):
        assert ra[g].status == 0 and ra[g].converged and rb[g].converged, g
        assert ra[g].iterations == rb[g].iterations and ra[g].reresid_inf == rb[g].reresid_inf, g
        assert np.array_equal(Pa[g].view(np.uint64), Pb[g].view(np.uint64)), g
        for x, y in ((wa[g].u, wb[g].u), (wa[g].v, wb[g].v), (wa[g].wv, wb[g].wv)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), g
