"""Parity of the 2D projection (shkz_project2d*, the macpressuresolver2 drop-in) with the 2D oracle.

Bars as in test_gpu_parity.py: identical unknown set, bit-identical couplings and rhs, diagonal
within 1e-14, pressure within 1e-5 rel-L2 of an FP64 oracle solve at a 1e-9 residual, velocities
bit-identical to the oracle's update applied to the GPU pressure, and a batch of 2D grids
bit-identical to the same grids projected alone.
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


SCENES = {
    "drop96": lambda: S.make_scene2d(S.SCENE_DROP, 96, seed=0),
    "solids80x64": lambda: S.make_scene2d(S.SCENE_DROP_SOLIDS, 80, 64, seed=1),
    "box70x90": lambda: S.make_scene2d(S.SCENE_SPHERES_BOX, 70, 90, seed=2),
    "odd": lambda: S.make_scene2d(S.SCENE_RANDOM_DROP, 37, 23, seed=9, fluid_dtype=S.F32),
}
OPERATOR_SCENES = {**SCENES, "smoke60x50": lambda: S.make_scene2d(S.SCENE_SMOKE, 60, 50, seed=3)}


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


@pytest.mark.parametrize("name", list(OPERATOR_SCENES))
def test_2d_operator_matches_oracle(ctx, name):
    say(f"2d operator {name}")
    sc = OPERATOR_SCENES[name]()
    exp = ctx.export_system2d(sc)
    A = O.assemble2(sc)
    im = A["index_map"]
    unk = im >= 0
    assert np.array_equal(exp["diag"] > 0, unk)
    n = A["n"]
    rows = np.repeat(np.arange(n), np.diff(A["rowptr"]))
    diag = np.zeros(n)
    off = {}
    for r_, c_, v_ in zip(rows, A["col"], A["val"]):
        if r_ == c_:
            diag[r_] = v_
        else:
            off[(r_, c_)] = v_
    gd = exp["diag"][unk][np.argsort(im[unk])]
    assert np.allclose(gd, diag, rtol=1e-14, atol=0)
    seen = 0
    for dim, key in enumerate(("gx", "gy")):
        g = exp[key]
        for (i, j) in np.argwhere(g != 0):
            q = (i - (dim == 0), j - (dim == 1))
            r1, r2 = im[i, j], im[q]
            assert r1 >= 0 and r2 >= 0
            assert -g[i, j] == off[(r1, r2)] == off[(r2, r1)]
            seen += 2
    assert seen == len(off)
    rhs = exp["rhs"][unk][np.argsort(im[unk])]
    assert np.array_equal(rhs, A["b"])


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED])
def test_2d_pressure_and_velocity(ctx, name, prec):
    say(f"2d mgpcg {name} prec={prec}")
    sc = SCENES[name]()
    A = O.assemble2(sc)
    x_ref, it, rr, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_APPLIED)
    assert ok
    work = sc.copy()
    r, P, P64 = ctx.project2d(work, S.default_params(residual=1e-9, precision=prec), pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    x_g = O.gather(A, P64)
    assert np.abs(A["b"] - O.spmv(A, x_g)).max() <= 1e-9 * np.abs(A["b"]).max() * 1.0001
    assert rel_l2(x_g, x_ref) < 1e-5
    ref = sc.copy()
    O.velocity_update2(ref, P, A["area"], A["rho"])
    assert np.array_equal(work.u, ref.u) and np.array_equal(work.v, ref.v)


def test_2d_reference_cg_iteration(ctx):
    say("2d reference cg")
    sc = SCENES["solids80x64"]()
    A = O.assemble2(sc)
    x_o, it_o, rr_o, ok = O.pcg(A, residual=1e-8, mic_mode=O.MIC_NONE)
    r, P, P64 = ctx.project2d(sc.copy(), S.default_params(residual=1e-8, solver=S.SOLVER_CG,
                                                          precision=S.PREC_FP64), pressure64=True)
    lo, hi = O.cg_count_envelope(A, residual=1e-8)  # the oracle's counts under other unknown orders
    assert r.converged and lo <= r.iterations <= hi, (r.iterations, it_o, lo, hi)


def test_2d_surface_tension_volume_correction_and_full_oracle(ctx):
    for kw in [dict(surface_tension=0.05), dict(second_order_fluid=0), dict(second_order_solid=0)]:
        say(f"2d options {kw}")
        sc = S.make_scene2d(S.SCENE_DROP_SOLIDS, 64, seed=7, fluid_dtype=S.F32)
        oprm = O.default_params(residual=1e-11, mic_mode=O.MIC_APPLIED)
        for k, v in kw.items():
            setattr(oprm, k, v)
        ref = sc.copy()
        r_o, P_o, P64_o = O.project2(ref, oprm)
        work = sc.copy()
        r, P, P64 = ctx.project2d(work, S.default_params(residual=1e-11, precision=S.PREC_FP64, **kw),
                                  pressure64=True)
        assert r.status == 0 and r.n_unknowns == r_o.n_unknowns
        assert rel_l2(P64, P64_o) < 1e-6
        assert np.abs(work.u - ref.u).max() <= 1e-4 * max(1.0, np.abs(ref.u).max())
        assert np.abs(work.v - ref.v).max() <= 1e-4 * max(1.0, np.abs(ref.v).max())


def test_2d_no_liquid_and_all_liquid(ctx):
    say("2d edge cases")
    w, h = 12, 9
    sc = S.Scene2D(w, h, 1 / 120, np.full((w, h), 0.5), None, *(np.ones(s, np.float32) for s in O.face_shapes2(w, h)))
    ref = sc.copy()
    O.project2(ref, O.default_params())
    work = sc.copy()
    r, P, _ = ctx.project2d(work)
    assert r.status == 0 and r.n_unknowns == 0 and not P.any()
    assert np.array_equal(work.u, ref.u) and np.array_equal(work.v, ref.v)
    for dims in [(1, 1), (2, 3), (8, 8), (17, 9)]:
        w, h = dims
        rng = np.random.default_rng(w * 31 + h)
        sc = S.Scene2D(w, h, 1 / 120, np.full(dims, -0.5), None,
                       *(rng.normal(size=s).astype(np.float32) for s in O.face_shapes2(w, h)))
        ref = sc.copy()
        O.project2(ref, O.default_params(residual=1e-10, mic_mode=O.MIC_APPLIED))
        work = sc.copy()
        r, P, _ = ctx.project2d(work, S.default_params(residual=1e-10, precision=S.PREC_FP64))
        assert r.status == 0
        for a, b in ((work.u, ref.u), (work.v, ref.v)):
            assert np.abs(a - b).max() <= 1e-4 * max(1.0, np.abs(b).max())


def test_2d_batch_equals_solo(ctx):
    say("2d batch vs solo")
    from shiokaze_amd.devmem import DeviceArray, Stream
    scenes = [S.make_scene2d(S.SCENE_RANDOM_DROP, n, m, seed=n + m) for n, m in ((40, 40), (8, 8), (64, 48), (5, 3))]
    prm = S.default_params(residual=1e-6)
    solo = []
    for sc in scenes:
        work = sc.copy()
        r, _, P64 = ctx.project2d(work, prm, pressure64=True)
        solo.append((r, work, P64))
    bufs, grids = [], []
    for sc in scenes:
        b = dict(fl=DeviceArray(sc.fluid), u=DeviceArray(sc.u), v=DeviceArray(sc.v),
                 p=DeviceArray(shape=(sc.w, sc.h), dtype=np.float64))
        bufs.append(b)
        grids.append(S.Grid2D(w=sc.w, h=sc.h, dx=0.0, dt=sc.dt, fluid=b["fl"].data_ptr(), fluid_dtype=S.F64,
                              solid_dtype=S.F32, solid=None, u=b["u"].data_ptr(), v=b["v"].data_ptr(),
                              pressure64=b["p"].data_ptr()))
    st = Stream()
    rb = ctx.project2d_batch_device(grids, prm, stream=st.handle)
    st.synchronize()
    for g, b in enumerate(bufs):
        r1, w1, P1 = solo[g]
        assert rb[g].iterations == r1.iterations and rb[g].reresid == r1.reresid
        assert np.array_equal(b["p"].download(), P1)
        assert np.array_equal(b["u"].download(), w1.u) and np.array_equal(b["v"].download(), w1.v)
        for a in b.values():
            a.free()
    st.destroy()


def test_2d_accuracy_known_answer(ctx):
    """accuracytest2 through the GPU path: same second-order error as the oracle."""
    say("2d accuracy KAT")
    import math
    errs = []
    for n in (32, 64):
        dx = 1.0 / n
        worst = 0.0
        for q in (-2, -1, 1, 2, 0):
            r0 = 0.25 + dx / 4 * q
            ii = (np.arange(n) + 0.5) * dx
            X, Y = np.meshgrid(ii, ii, indexing="ij")
            fluid = (np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - r0).astype(np.float32)
            fi = np.arange(n + 1) * dx
            u = np.broadcast_to(2.0 * (fi - 0.5)[:, None], (n + 1, n)).astype(np.float32).copy()
            v = np.broadcast_to(2.0 * (fi - 0.5)[None, :], (n, n + 1)).astype(np.float32).copy()
            sc = S.Scene2D(n, n, 1.0, np.ascontiguousarray(fluid), None, u, v)
            r, P, _ = ctx.project2d(sc, S.default_params(residual=1e-12, eps_fluid=1e-18, precision=S.PREC_FP64))
            exact = (X - 0.5) ** 2 + (Y - 0.5) ** 2 - r0 * r0
            mask = P != 0
            worst = max(worst, float(np.abs(exact[mask] - P[mask]).max()))
        errs.append(worst)
    order = math.log2(errs[0] / errs[1])
    say(f"  errors {errs} order {order:.2f}")
    assert errs[1] < 1e-3 and order > 1.6
