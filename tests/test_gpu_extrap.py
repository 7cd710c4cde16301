"""The step after the projection on the GPU (shkz_extrapolate_velocity*, the reference's
macutility3::extrapolate_and_constrain_velocity) against the oracle, bit for bit:

  * the post-projection face activity (shkz_grid.active_out_*) equals the oracle's velocity update
    rule (macpressuresolver3.cpp:252-266);
  * extrapolation for widths 1..6 and the solid constraint, on the projected velocities of scenes
    with solid spheres, a solid box and an all-fluid smoke box, a float and a double solid level
    set, odd shapes: velocities and activity identical to the oracle's;
  * the constraint alone (width 0), no solid, a solid without a negative value;
  * the device entry right after shkz_project_device on one stream, with no host synchronisation,
    equals the host entry;
  * invalid input is refused.
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
    "solids40": lambda: S.make_scene(S.SCENE_DROP_SOLIDS, 40, seed=1),
    "box36": lambda: S.make_scene(S.SCENE_SPHERES_BOX, 36, seed=2),
    "smoke32": lambda: S.make_scene(S.SCENE_SMOKE, 32, seed=3),
    "odd": lambda: S.make_scene(S.SCENE_DROP_SOLIDS, 23, 17, 30, seed=4, fluid_dtype=S.F32),
}


def projected(ctx, sc):
    """Projects sc on the GPU; returns (scene after projection, GPU activity, oracle activity)."""
    work = sc.copy()
    shapes = O.face_shapes(sc.w, sc.h, sc.d)
    act = [np.full(s, 7, np.uint8) for s in shapes]
    r, P, _ = ctx.project(work, S.default_params(residual=1e-6), active_out=act)
    assert r.status == 0 and r.converged
    area, rho = O.fractions(sc)
    ref = sc.copy()
    act_o = O.velocity_update(ref, P, area, rho, active_out=True)
    return work, act, act_o, ref


@pytest.mark.parametrize("name", list(SCENES))
def test_post_projection_activity_matches_oracle(ctx, name):
    say(f"active_out {name}")
    sc = SCENES[name]()
    work, act, act_o, ref = projected(ctx, sc)
    for a in range(3):
        assert np.array_equal(act[a], act_o[a]), a
        assert act[a].max() <= 1 and act[a].any()
    for x, y in ((work.u, ref.u), (work.v, ref.v), (work.wv, ref.wv)):
        assert np.array_equal(x, y)


def run_both(ctx, sc, vel, act, width, constrain, solid):
    vg = [v.copy() for v in vel]
    ag = [a.copy() for a in act]
    ctx.extrapolate_velocity(sc.w, sc.h, sc.d, vg, ag, solid=solid, width=width, constrain=constrain)
    vo = [v.copy() for v in vel]
    ao = [a.copy() for a in act]
    O.extrapolate(sc.w, sc.h, sc.d, vo, ao, width)
    if constrain and solid is not None:
        O.constrain_velocity(sc.w, sc.h, sc.d, vo, ao, solid)
    return vg, ag, vo, ao


def assert_same(vg, ag, vo, ao):
    for a in range(3):
        assert np.array_equal(ag[a], ao[a]), f"activity {a}"
        diff = np.flatnonzero(vg[a].view(np.uint32) != vo[a].view(np.uint32))
        assert diff.size == 0, (a, diff[:5], vg[a].ravel()[diff[:5]], vo[a].ravel()[diff[:5]])


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("width", [1, 3, 6])
def test_extrapolate_and_constrain_match_oracle(ctx, name, width):
    say(f"extrapolate {name} width={width}")
    sc = SCENES[name]()
    work, act, _, _ = projected(ctx, sc)
    vel = [work.u, work.v, work.wv]
    solid = sc.solid
    assert solid is not None
    vg, ag, vo, ao = run_both(ctx, sc, vel, act, width, True, solid)
    assert_same(vg, ag, vo, ao)
    grew = sum(int(a.sum()) for a in ag) - sum(int(a.sum()) for a in act)
    say(f"  faces activated {grew}")
    assert grew > 0


def test_double_solid_and_options(ctx):
    sc = SCENES["solids40"]()
    work, act, _, _ = projected(ctx, sc)
    vel = [work.u, work.v, work.wv]
    say("extrapolate double solid")
    vg, ag, vo, ao = run_both(ctx, sc, vel, act, 2, True, sc.solid.astype(np.float64))
    assert_same(vg, ag, vo, ao)
    say("constraint only")
    vg, ag, vo, ao = run_both(ctx, sc, vel, act, 0, True, sc.solid)
    assert_same(vg, ag, vo, ao)
    assert all(np.array_equal(x, y) for x, y in zip(ag, act))
    say("extrapolation only")
    vg, ag, vo, ao = run_both(ctx, sc, vel, act, 4, False, sc.solid)
    assert_same(vg, ag, vo, ao)
    say("solid without a negative value")
    pos = np.abs(sc.solid) + 0.01
    vg, ag, vo, ao = run_both(ctx, sc, vel, act, 2, True, pos)
    assert_same(vg, ag, vo, ao)
    vg2, ag2, _, _ = run_both(ctx, sc, vel, act, 2, False, None)
    assert all(np.array_equal(x, y) for x, y in zip(vg, vg2))


def test_random_activity_and_values(ctx):
    """Scattered activity on random values (no projection): every pass and the constraint."""
    say("extrapolate random field")
    rng = np.random.default_rng(12)
    sc = S.make_scene(S.SCENE_SPHERES_BOX, 29, 31, 27, seed=6)
    shapes = O.face_shapes(sc.w, sc.h, sc.d)
    vel = [rng.normal(size=s).astype(np.float32) for s in shapes]
    act = [(rng.uniform(size=s) < 0.05).astype(np.uint8) for s in shapes]
    for width in (1, 2, 5):
        vg, ag, vo, ao = run_both(ctx, sc, vel, act, width, True, sc.solid)
        assert_same(vg, ag, vo, ao)


def test_device_entry_follows_projection_on_one_stream(ctx):
    say("extrapolate device entry after project_device")
    from shiokaze_amd.devmem import DeviceArray, Stream
    sc = SCENES["box36"]()
    shapes = O.face_shapes(sc.w, sc.h, sc.d)
    fl, so = DeviceArray(sc.fluid), DeviceArray(sc.solid)
    vel = [DeviceArray(a) for a in (sc.u, sc.v, sc.wv)]
    act = [DeviceArray(np.zeros(s, np.uint8)) for s in shapes]
    st = Stream()
    prm = S.default_params(residual=1e-6)
    r = ctx.project_device(sc.w, sc.h, sc.d, sc.dt, fl, S.F64, so, S.F32, *vel, params=prm, stream=st.handle,
                           active_out=act)
    ctx.extrapolate_velocity_device(sc.w, sc.h, sc.d, vel, act, solid=so, solid_dtype=S.F32, width=3,
                                    stream=st.handle)
    st.synchronize()
    got_v = [v.download() for v in vel]
    got_a = [a.download() for a in act]
    st.destroy()
    for a in [fl, so] + vel + act:
        a.free()
    assert r.status == 0
    work, act_h, _, _ = projected(ctx, sc)
    vh = [work.u, work.v, work.wv]
    ctx.extrapolate_velocity(sc.w, sc.h, sc.d, vh, act_h, solid=sc.solid, width=3)
    for a in range(3):
        assert np.array_equal(got_a[a], act_h[a])
        assert np.array_equal(got_v[a], vh[a])


def test_invalid_input_is_refused(ctx):
    say("extrapolate invalid input")
    w, h, d = 6, 5, 1
    shapes = O.face_shapes(w, h, d)
    vel = [np.zeros(s, np.float32) for s in shapes]
    act = [np.ones(s, np.uint8) for s in shapes]
    solid = np.full((w + 1, h + 1, d + 1), -1.0, np.float32)
    assert ctx.extrapolate_velocity(w, h, d, vel, act, solid=solid, width=1, check=False) == S.ERR_INVALID
    assert ctx.extrapolate_velocity(w, h, d, vel, act, solid=None, width=300, check=False) == S.ERR_INVALID
    # without the constraint a one-cell-thick grid is fine
    assert ctx.extrapolate_velocity(w, h, d, vel, act, solid=solid, width=2, constrain=False) == 0
