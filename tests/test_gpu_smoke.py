"""The all-fluid (smoke) branch of the projection: a fluid level set below zero everywhere makes
every cell with an open face an unknown and rho = 1 on every face (macpressuresolver3.cpp:148-152,
called from macsmoke3.cpp:122,293 with m_fluid initialised to -1, macsmoke3.cpp:106).  With
closed walls the system is pure Neumann: singular, with the constants of each connected component
as its null space, and consistent (the right-hand side sums to zero per component).

Bars: the unknown set and the operator as in test_gpu_parity.py (checked there on the smoke
scenes too); pressure equal to the FP64 oracle's up to one constant per connected component
(rel-L2 1e-5 after removing each component's mean); the FP64 residual of the returned pressure
meets the requested tolerance; velocities bit-identical to the oracle's update applied to the
GPU pressure; the divergence gate of test_gpu_parity.py; the reference CG iteration's count
within the envelope of the oracle's counts under other unknown orders (rounding alone)."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

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


def demean(A, x):
    """x minus its mean over each connected component of the operator's graph."""
    n = A["n"]
    M = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(n, n))
    nc, lab = connected_components(M, directed=False)
    cnt = np.bincount(lab, minlength=nc)
    mean = np.bincount(lab, weights=x, minlength=nc) / np.maximum(cnt, 1)
    return x - mean[lab], nc


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


SCENES3 = {
    "smoke48": lambda: S.make_scene(S.SCENE_SMOKE, 48, seed=3),
    "smoke_odd": lambda: S.make_scene(S.SCENE_SMOKE, 37, 21, 30, seed=4, fluid_dtype=S.F32),
}


@pytest.mark.parametrize("name", list(SCENES3))
@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED])
def test_smoke_pressure_and_velocity(ctx, name, prec):
    say(f"smoke mgpcg {name} prec={prec}")
    sc = SCENES3[name]()
    A = O.assemble(sc)
    x_ref, it, rr, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_APPLIED)
    assert ok
    work = sc.copy()
    r, P, P64 = ctx.project(work, S.default_params(residual=1e-9, precision=prec), pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    assert r.n_unknowns == A["n"]
    x_g = O.gather(A, P64)
    assert np.abs(A["b"] - O.spmv(A, x_g)).max() <= 1e-9 * np.abs(A["b"]).max() * 1.0001
    xg0, nc = demean(A, x_g)
    xr0, _ = demean(A, x_ref)
    say(f"  components {nc}, iterations {r.iterations}")
    assert rel_l2(xg0, xr0) < 1e-5
    ref = sc.copy()
    O.velocity_update(ref, P, A["area"], A["rho"])
    for a, b in ((work.u, ref.u), (work.v, ref.v), (work.wv, ref.wv)):
        assert np.array_equal(a, b)


def test_smoke_divergence_gate(ctx):
    say("smoke divergence")
    sc = SCENES3["smoke48"]()
    A = O.assemble(sc)
    d0, _ = O.divergence(sc, A["index_map"], A["area"])
    ref = sc.copy()
    O.project(ref, O.default_params(residual=1e-10, mic_mode=O.MIC_APPLIED))
    work = sc.copy()
    r, P, _ = ctx.project(work, S.default_params(residual=1e-6))
    assert r.status == 0 and r.converged
    d_gpu, _ = O.divergence(work, A["index_map"], A["area"])
    d_ref, _ = O.divergence(ref, A["index_map"], A["area"])
    say(f"  max div before {d0:.3e}  gpu {d_gpu:.3e}  oracle {d_ref:.3e}")
    assert abs(d_gpu - d_ref) / d0 <= 1e-5


def test_smoke_reference_cg_iteration(ctx):
    say("smoke reference cg")
    sc = S.make_scene(S.SCENE_SMOKE, 32, seed=3)
    A = O.assemble(sc)
    x_o, it_o, rr_o, ok = O.pcg(A, residual=1e-8, mic_mode=O.MIC_NONE)
    assert ok
    r, P, P64 = ctx.project(sc.copy(), S.default_params(residual=1e-8, solver=S.SOLVER_CG, precision=S.PREC_FP64),
                            pressure64=True)
    lo, hi = O.cg_count_envelope(A, residual=1e-8)  # the oracle's counts under other unknown orders
    assert r.converged and lo <= r.iterations <= hi, (r.iterations, it_o, lo, hi)


@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED])
def test_smoke_2d(ctx, prec):
    say(f"smoke 2d prec={prec}")
    sc = S.make_scene2d(S.SCENE_SMOKE, 90, 70, seed=5)
    assert (sc.fluid == -1.0).all()
    A = O.assemble2(sc)
    x_ref, it, rr, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_APPLIED)
    assert ok
    work = sc.copy()
    r, P, P64 = ctx.project2d(work, S.default_params(residual=1e-9, precision=prec), pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    x_g = O.gather(A, P64)
    assert np.abs(A["b"] - O.spmv(A, x_g)).max() <= 1e-9 * np.abs(A["b"]).max() * 1.0001
    assert rel_l2(demean(A, x_g)[0], demean(A, x_ref)[0]) < 1e-5
    ref = sc.copy()
    O.velocity_update2(ref, P, A["area"], A["rho"])
    assert np.array_equal(work.u, ref.u) and np.array_equal(work.v, ref.v)


def test_smoke_batch_with_liquid_scenes(ctx):
    """Singular smoke grids and liquid grids in one batch: each result equals its solo solve."""
    say("smoke batch")
    scenes = [S.make_scene(S.SCENE_SMOKE, 24, seed=1), S.make_scene(S.SCENE_DROP, 32, seed=2),
              S.make_scene(S.SCENE_SMOKE, 19, 26, 17, seed=3)]
    prm = S.default_params(residual=1e-7)
    solo = []
    for sc in scenes:
        w = sc.copy()
        r, P, _ = ctx.project(w, prm)
        solo.append((r, P, w))
    works = [sc.copy() for sc in scenes]
    rb, Pb, _ = ctx.project_batch(works, prm)
    for (r0, P0, w0), r1, P1, w1 in zip(solo, rb, Pb, works):
        assert r1.status == 0 and r1.converged and r1.iterations == r0.iterations
        assert np.array_equal(P1, P0)
        for a, b in ((w1.u, w0.u), (w1.v, w0.v), (w1.wv, w0.wv)):
            assert np.array_equal(a, b)
