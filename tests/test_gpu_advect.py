"""Advection on the GPU (shkz_advect_velocity*, shkz_advect_scalar*: the reference's
macadvection3 with its defaults) against the oracle, bit for bit:

  * the velocity a projection and the extrapolation leave, advected by itself (MacCormack and
    semi-Lagrangian, TrimNarrowBand 1 and 2, float and double liquid level sets);
  * the liquid level set advected by that velocity with the narrow band measured on itself (the
    level-set tracker's call), on a narrow band of active cells with a background value, and a
    scalar with an explicit level set;
  * the WENO option (six-point WENO sampling in the value passes) on the velocity and on the level
    set, MacCormack and semi-Lagrangian, with a narrow band of active cells;
  * the whole step chained on the device on one stream (projection, extrapolation and
    constraint, velocity and level-set advection) equals the host entries;
  * invalid input is refused.
"""
import ctypes
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


def stepped(ctx, kind, n, seed, width=3):
    """A scene after a GPU projection and extrapolation: (scene, velocities, activity)."""
    sc = S.make_scene(kind, n, seed=seed)
    work = sc.copy()
    act = [np.zeros(s, np.uint8) for s in O.face_shapes(sc.w, sc.h, sc.d)]
    r, P, _ = ctx.project(work, S.default_params(residual=1e-6), active_out=act)
    assert r.status == 0
    vel = [work.u, work.v, work.wv]
    ctx.extrapolate_velocity(sc.w, sc.h, sc.d, vel, act, solid=sc.solid, width=width)
    return sc, vel, act


def same(a, b):
    d = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
    return d.size == 0, (d[:5], a.ravel()[d[:5]], b.ravel()[d[:5]])


CASES = [(S.SCENE_DROP_SOLIDS, 40, 1), (S.SCENE_SPHERES_BOX, 36, 2), (S.SCENE_RANDOM_DROP, 33, 9)]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("mc,trim", [(True, 1), (False, 1), (True, 2)])
def test_velocity_advection_matches_oracle(ctx, case, mc, trim):
    kind, n, seed = CASES[case]
    say(f"advect velocity case={case} mc={mc} trim={trim}")
    sc, vel, act = stepped(ctx, kind, n, seed)
    dt = 2.0 * sc.dx / max(float(np.abs(v).max()) for v in vel)  # CFL 2
    fluid = sc.fluid.astype(np.float32)
    vg = [v.copy() for v in vel]
    ctx.advect_velocity(sc.w, sc.h, sc.d, vg, act, fluid, dt, maccormack=mc, trim=trim)
    vo = [v.copy() for v in vel]
    O.advect_velocity(sc.w, sc.h, sc.d, vo, act, fluid, dt, maccormack=mc, trim=trim)
    for a in range(3):
        ok, info = same(vg[a], vo[a])
        assert ok, (a, info)
    assert any(not np.array_equal(x, y) for x, y in zip(vg, vel))


def test_velocity_advection_double_level_set(ctx):
    say("advect velocity f64 level set")
    sc, vel, act = stepped(ctx, S.SCENE_DROP_SOLIDS, 32, 3)
    dt = 1.5 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    vg = [v.copy() for v in vel]
    ctx.advect_velocity(sc.w, sc.h, sc.d, vg, act, sc.fluid, dt)
    vo = [v.copy() for v in vel]
    O.advect_velocity(sc.w, sc.h, sc.d, vo, act, sc.fluid, dt)
    for a in range(3):
        assert same(vg[a], vo[a])[0], a


@pytest.mark.parametrize("mc", [True, False])
def test_level_set_advection_matches_oracle(ctx, mc):
    say(f"advect level set mc={mc}")
    sc, vel, act = stepped(ctx, S.SCENE_SPHERES_BOX, 40, 2)
    dt = 2.0 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    phi = sc.fluid.astype(np.float32)
    # whole grid, narrow band measured on phi itself
    qg, qo = phi.copy(), phi.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, vel, act, qg, dt, maccormack=mc)
    O.advect_scalar(sc.w, sc.h, sc.d, vel, act, qo, dt, maccormack=mc)
    assert same(qg, qo)[0]
    assert not np.array_equal(qg, phi)
    # a narrow band of active cells (|phi| < 3 dx) over a background value
    band = (np.abs(phi) < 3 * sc.dx).astype(np.uint8)
    qg, qo = phi.copy(), phi.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, vel, act, qg, dt, active_q=band, maccormack=mc, background=0.125)
    O.advect_scalar(sc.w, sc.h, sc.d, vel, act, qo, dt, active_q=band, maccormack=mc, background=0.125)
    assert same(qg, qo)[0]
    assert np.array_equal(qg[band == 0], phi[band == 0])


def test_scalar_with_explicit_level_set(ctx):
    say("advect scalar with level set")
    sc, vel, act = stepped(ctx, S.SCENE_DROP, 30, 0)
    rng = np.random.default_rng(2)
    q = rng.normal(size=(sc.w, sc.h, sc.d)).astype(np.float32)
    dt = 1.0 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    for fl in (sc.fluid.astype(np.float32), sc.fluid):
        qg, qo = q.copy(), q.copy()
        ctx.advect_scalar(sc.w, sc.h, sc.d, vel, act, qg, dt, fluid=fl, trim=2)
        O.advect_scalar(sc.w, sc.h, sc.d, vel, act, qo, dt, fluid=fl, trim=2)
        assert same(qg, qo)[0]


@pytest.mark.parametrize("case", [0, 2])
@pytest.mark.parametrize("mc", [True, False])
def test_weno_velocity_advection_matches_oracle(ctx, case, mc):
    kind, n, seed = CASES[case]
    say(f"advect velocity WENO case={case} mc={mc}")
    sc, vel, act = stepped(ctx, kind, n, seed)
    dt = 1.5 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    fluid = sc.fluid.astype(np.float32)
    vg = [v.copy() for v in vel]
    ctx.advect_velocity(sc.w, sc.h, sc.d, vg, act, fluid, dt, maccormack=mc, weno=True)
    vo = [v.copy() for v in vel]
    O.advect_velocity(sc.w, sc.h, sc.d, vo, act, fluid, dt, maccormack=mc, weno=True)
    for a in range(3):
        ok, info = same(vg[a], vo[a])
        assert ok, (a, info)
    vt = [v.copy() for v in vel]
    ctx.advect_velocity(sc.w, sc.h, sc.d, vt, act, fluid, dt, maccormack=mc)
    assert any(not np.array_equal(x, y) for x, y in zip(vg, vt))


@pytest.mark.parametrize("mc", [True, False])
def test_weno_level_set_advection_matches_oracle(ctx, mc):
    say(f"advect level set WENO mc={mc}")
    sc, vel, act = stepped(ctx, S.SCENE_SPHERES_BOX, 36, 2)
    dt = 2.0 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    phi = sc.fluid.astype(np.float32)
    qg, qo = phi.copy(), phi.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, vel, act, qg, dt, maccormack=mc, weno=True)
    O.advect_scalar(sc.w, sc.h, sc.d, vel, act, qo, dt, maccormack=mc, weno=True)
    assert same(qg, qo)[0]
    band = (np.abs(phi) < 3 * sc.dx).astype(np.uint8)
    qg, qo = phi.copy(), phi.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, vel, act, qg, dt, active_q=band, maccormack=mc, background=0.125, weno=True)
    O.advect_scalar(sc.w, sc.h, sc.d, vel, act, qo, dt, active_q=band, maccormack=mc, background=0.125, weno=True)
    assert same(qg, qo)[0]
    assert np.array_equal(qg[band == 0], phi[band == 0])


def test_weno_on_device_arrays_matches_host_entry(ctx):
    say("advect WENO device entries")
    from shiokaze_amd.devmem import DeviceArray, Stream
    sc, vel, act = stepped(ctx, S.SCENE_DROP_SOLIDS, 32, 4)
    dt = 1.0 * sc.dx / max(float(np.abs(v).max()) for v in vel)
    phi = sc.fluid.astype(np.float32)
    dv = [DeviceArray(v) for v in vel]
    da = [DeviceArray(a) for a in act]
    dphi = DeviceArray(phi)
    st = Stream()
    ctx.advect_scalar_device(sc.w, sc.h, sc.d, dv, da, dphi, dt, stream=st.handle, weno=True)
    ctx.advect_velocity_device(sc.w, sc.h, sc.d, dv, da, dphi, S.F32, dt, stream=st.handle, weno=True)
    st.synchronize()
    gv = [v.download() for v in dv]
    gphi = dphi.download()
    st.destroy()
    for a in dv + da + [dphi]:
        a.free()
    hv = [v.copy() for v in vel]
    hphi = phi.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, hv, act, hphi, dt, weno=True)
    ctx.advect_velocity(sc.w, sc.h, sc.d, hv, act, hphi, dt, weno=True)
    assert np.array_equal(gphi, hphi)
    for a in range(3):
        assert np.array_equal(gv[a], hv[a])


def test_whole_step_chained_on_device(ctx):
    """Projection, extrapolation and constraint, velocity and level-set advection on device
    arrays, queued on one stream without host synchronisation, equal the host entries."""
    say("whole step on device")
    from shiokaze_amd.devmem import DeviceArray, Stream
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 40, seed=5, fluid_dtype=S.F32)
    shapes = O.face_shapes(sc.w, sc.h, sc.d)
    dt = 0.004
    fl, so = DeviceArray(sc.fluid), DeviceArray(sc.solid)
    vel = [DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]
    act = [DeviceArray(np.zeros(s, np.uint8)) for s in shapes]
    st = Stream()
    ctx.project_device(sc.w, sc.h, sc.d, sc.dt, fl, S.F32, so, S.F32, *vel, params=S.default_params(residual=1e-6),
                       stream=st.handle, active_out=act)
    ctx.extrapolate_velocity_device(sc.w, sc.h, sc.d, vel, act, solid=so, width=3, stream=st.handle)
    ctx.advect_scalar_device(sc.w, sc.h, sc.d, vel, act, fl, dt, stream=st.handle)
    ctx.advect_velocity_device(sc.w, sc.h, sc.d, vel, act, fl, S.F32, dt, stream=st.handle)
    st.synchronize()
    gv = [v.download() for v in vel]
    gphi = fl.download()
    st.destroy()
    for a in [fl, so] + vel + act:
        a.free()
    # host path, step by step
    work = sc.copy()
    ha = [np.zeros(s, np.uint8) for s in shapes]
    ctx.project(work, S.default_params(residual=1e-6), active_out=ha)
    hv = [work.u, work.v, work.wv]
    ctx.extrapolate_velocity(sc.w, sc.h, sc.d, hv, ha, solid=sc.solid, width=3)
    phi = sc.fluid.copy()
    ctx.advect_scalar(sc.w, sc.h, sc.d, hv, ha, phi, dt)
    ctx.advect_velocity(sc.w, sc.h, sc.d, hv, ha, phi, dt)
    assert np.array_equal(gphi, phi)
    for a in range(3):
        assert np.array_equal(gv[a], hv[a])


def test_invalid_input_is_refused(ctx):
    say("advect invalid input")
    w, h, d = 6, 5, 1
    shapes = O.face_shapes(w, h, d)
    vel = [np.zeros(s, np.float32) for s in shapes]
    act = [np.ones(s, np.uint8) for s in shapes]
    q = np.zeros((w, h, d), np.float32)
    assert ctx.advect_scalar(w, h, d, vel, act, q, 0.1, check=False) == S.ERR_INVALID
    w, h, d = 6, 5, 4
    shapes = O.face_shapes(w, h, d)
    vel = [np.zeros(s, np.float32) for s in shapes]
    act = [np.ones(s, np.uint8) for s in shapes]
    assert ctx.advect_velocity(w, h, d, vel, act, None, 0.1, check=False) == S.ERR_INVALID
    assert ctx.advect_velocity(w, h, d, vel, act, np.zeros((w, h, d), np.float32), float("nan"), check=False) == \
        S.ERR_INVALID
    q = np.zeros((w, h, d), np.float32)
    p = S.AdvectParams()
    S.lib().shkz_advect_default_params(ctypes.byref(p))
    p.dt, p.weno = 0.1, 2  # the flag is 0 or 1
    p.fluid, p.fluid_dtype = None, S.F32
    f = ctx._vfield(w, h, d, vel, act, 0.0, False)
    assert S.lib().shkz_advect_scalar(ctx.handle, ctypes.byref(f), q.ctypes.data, None, ctypes.byref(p)) == \
        S.ERR_INVALID
    # the kernels index face, cell and node arrays in 32 bits: a grid with (w+1)(h+1)(d+1) >= 2^32
    # (1701^3 = 4.9e9) is refused before any array is read, so the small arrays stand in for it
    big = ctx._vfield(1700, 1700, 1700, vel, act, 0.0, False)
    p.weno = 0
    assert S.lib().shkz_advect_velocity(ctx.handle, ctypes.byref(big), ctypes.byref(p)) == S.ERR_INVALID
    assert "32-bit" in S.last_error()
