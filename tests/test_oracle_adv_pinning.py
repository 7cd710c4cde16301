"""Pins the advection oracle (oracle/shkz_oracle_adv.c, macadvection3) without a GPU.  The
reference's advection needs its sparse-array module system and is not buildable here; its
interpolation weights are pinned bit for bit elsewhere (test_oracle_vel_pinning.py), and the
rest by closed forms chosen so that a wrong sign, index, component or pass order fails:

  * zero velocity changes nothing; a uniform field advected by itself with a dyadic CFL stays
    exactly uniform;
  * a linear scalar in a uniform flow moves by exactly u dt / dx (semi-Lagrangian and MacCormack,
    whose correction vanishes on linear fields);
  * a face velocity field u = j (v = c) moves by the 4-face average of the other component
    (convert_to_full), exactly on dyadic values;
  * MacCormack stays within the min/max of its forward stencil and switches to the
    semi-Lagrangian value inside the narrow band (fluid > -dx * TrimNarrowBand);
  * inactive cells and faces are never written; cells whose full velocity is incomplete keep
    their value.
"""
import numpy as np
import pytest

from oracle import oracle as O


def faces(w, h, d, fill=0.0, active=1):
    shapes = O.face_shapes(w, h, d)
    return [np.full(s, fill, np.float32) for s in shapes], [np.full(s, active, np.uint8) for s in shapes]


def deep_fluid(w, h, d):
    return np.full((w, h, d), -1.0, np.float32)


@pytest.mark.parametrize("mc", [False, True])
def test_zero_velocity_changes_nothing(mc):
    w, h, d = 9, 8, 7
    rng = np.random.default_rng(1)
    vel, act = faces(w, h, d)
    q = rng.normal(size=(w, h, d)).astype(np.float32)
    q0 = q.copy()
    O.advect_scalar(w, h, d, vel, act, q, 0.1, fluid=deep_fluid(w, h, d), maccormack=mc)
    assert np.array_equal(q, q0)
    O.advect_velocity(w, h, d, vel, act, deep_fluid(w, h, d), 0.1, maccormack=mc)
    assert all((v == 0).all() for v in vel)


@pytest.mark.parametrize("mc", [False, True])
def test_uniform_flow_stays_uniform(mc):
    w = h = d = 10
    dx = 0.1
    vel, act = faces(w, h, d)
    for a, c in enumerate((0.5, -0.25, 0.75)):
        vel[a][...] = c
    dt = 0.05  # |u| dt / dx = 0.25, 0.125, 0.375: dyadic weights
    O.advect_velocity(w, h, d, vel, act, deep_fluid(w, h, d), dt, dx=dx, maccormack=mc)
    for a, c in enumerate((0.5, -0.25, 0.75)):
        assert (vel[a] == np.float32(c)).all(), a


@pytest.mark.parametrize("mc", [False, True])
def test_linear_scalar_moves_exactly(mc):
    w, h, d = 16, 12, 10
    dx = 1.0 / 16
    vel, act = faces(w, h, d)
    vel[0][...] = 0.5
    vel[1][...] = -0.25
    dt = 0.0625  # shift (0.5, -0.25, 0) cells
    i, j, k = np.meshgrid(np.arange(w), np.arange(h), np.arange(d), indexing="ij")
    q = (2.0 * i + 4.0 * j + 1.0 * k).astype(np.float32)
    O.advect_scalar(w, h, d, vel, act, q, dt, fluid=deep_fluid(w, h, d), dx=dx, maccormack=mc)
    want = 2.0 * (i - 0.5) + 4.0 * (j + 0.25) + k
    inner = (i >= 2) & (i <= w - 3) & (j >= 2) & (j <= h - 3)
    assert np.array_equal(q[inner], want[inner].astype(np.float32))


def test_face_velocity_uses_four_face_average():
    w = h = d = 12
    dx = 1.0 / 12
    vel, act = faces(w, h, d)
    su, sv, sw = O.face_shapes(w, h, d)
    jj = np.arange(su[1])[None, :, None]
    vel[0][...] = np.broadcast_to(jj.astype(np.float32), su)
    vel[1][...] = 0.5
    dt = dx / 4  # v shifts u by 0.125 cells in y; u shifts v faces in x (v is constant in x)
    O.advect_velocity(w, h, d, vel, act, deep_fluid(w, h, d), dt, dx=dx, maccormack=False)
    # u faces: departure y = j - 0.5 * 0.25 (v averaged from 4 faces = 0.5); u constant in x and z
    inner = np.zeros(su, bool)
    inner[:, 2:h - 2, :] = True
    want = np.broadcast_to((jj - 0.125).astype(np.float32), su)
    assert np.array_equal(vel[0][inner], want[inner])
    assert (vel[1] == np.float32(0.5)).all() and (vel[2] == 0).all()


def test_maccormack_is_bounded_and_narrow_band_uses_semi_lagrangian():
    w = h = d = 14
    dx = 1.0 / 14
    rng = np.random.default_rng(4)
    vel, act = faces(w, h, d)
    for a in range(3):
        vel[a][...] = rng.uniform(-1, 1, size=vel[a].shape)
    q = (rng.uniform(size=(w, h, d)) > 0.5).astype(np.float32)  # a rough field
    fluid = np.where(np.arange(w)[:, None, None] < 7, -1.0, 0.5).astype(np.float32) * np.ones((w, h, d), np.float32)
    dt = 0.5 * dx
    q_sl, q_mc = q.copy(), q.copy()
    O.advect_scalar(w, h, d, vel, act, q_sl, dt, fluid=fluid, dx=dx, maccormack=False)
    O.advect_scalar(w, h, d, vel, act, q_mc, dt, fluid=fluid, dx=dx, maccormack=True)
    assert (q_mc >= 0).all() and (q_mc <= 1).all()          # bounded by the stencil min/max
    # departure points move at most 0.5 cell per axis: from x >= 9 they sample fluid = 0.5 only
    assert np.array_equal(q_mc[9:], q_sl[9:])               # narrow band: the semi-Lagrangian value
    assert not np.array_equal(q_mc[1:6], q_sl[1:6])         # MacCormack acts in the deep liquid


def test_inactive_entries_and_incomplete_velocity():
    w, h, d = 8, 8, 8
    dx = 1.0 / 8
    rng = np.random.default_rng(7)
    vel, act = faces(w, h, d)
    for a in range(3):
        vel[a][...] = rng.uniform(-1, 1, size=vel[a].shape)
    act[0][4, 3, 3] = 0                                    # cells (3,3,3) and (4,3,3) lose a face
    q = rng.normal(size=(w, h, d)).astype(np.float32)
    aq = np.ones((w, h, d), np.uint8)
    aq[0] = 0
    q0 = q.copy()
    O.advect_scalar(w, h, d, vel, act, q, 0.3 * dx, active_q=aq, fluid=deep_fluid(w, h, d), dx=dx)
    assert np.array_equal(q[0], q0[0])
    assert q[3, 3, 3] == q0[3, 3, 3] and q[4, 3, 3] == q0[4, 3, 3]
    v0 = [v.copy() for v in vel]
    O.advect_velocity(w, h, d, vel, act, deep_fluid(w, h, d), 0.3 * dx, dx=dx)
    assert vel[0][4, 3, 3] == v0[0][4, 3, 3]
