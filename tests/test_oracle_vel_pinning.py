"""Pins the oracle of the post-projection step (velocity extrapolation and the solid constraint,
oracle/shkz_oracle_vel.c) without a GPU:

  * interpolation and derivative weights: bit-identical to the reference's own header code
    (array_interpolator3::interpolate_coef, array_derivative3::derivative_interpolate_coef,
    compiled where they lie by oracle/Makefile) through the committed fixtures, and against the
    live reference build where it exists;
  * extrapolation (array_extrapolator3.h:47-80 over dilate3.h:43-101, which needs the module
    system and is not buildable here): the active set after `width` passes equals the 6-neighbour
    dilation of the initial set (scipy), passes read only the previous pass (two sources four faces
    apart meet after two passes, not one), a half-space extends by exact copies, means of equal
    neighbours and of 1 and 2 are exact;
  * constraint (macutility3.cpp:61-88): a planar solid removes exactly the normal component of an
    incoming velocity and keeps a leaving one, a tilted plane gives the tangential projection, the
    boundary rule zeroes outflow only where a solid exists, inactive faces are never touched.
"""
import os

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_interp_vectors.npz")


def test_interp_and_deriv_weights_match_reference_fixtures():
    z = np.load(GOLDEN)
    for q in range(len(z["shapes"])):
        idx, c = O.interp_coef(z["shapes"][q], z["pos"][q])
        assert np.array_equal(idx, z["interp_idx"][q]) and np.array_equal(c, z["interp_coef"][q]), q
        idx, c = O.deriv_coef(z["shapes"][q], z["pos"][q])
        assert np.array_equal(idx, z["deriv_idx"][q]) and np.array_equal(c, z["deriv_coef"][q]), q


@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="reference header driver not built")
def test_interp_and_deriv_weights_match_live_reference():
    rng = np.random.default_rng(11)
    for _ in range(500):
        sh = rng.integers(2, 30, size=3)
        p = rng.uniform(-2, sh + 1)
        a, b = O.interp_coef(sh, p), O.ref_interp_coef(sh, p)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        a, b = O.deriv_coef(sh, p), O.ref_deriv_coef(sh, p)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_weights_reproduce_trilinear_fields():
    """Interpolation weights are exact for trilinear functions, derivative weights give their
    gradient (closed forms)."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        sh = rng.integers(2, 12, size=3)
        c = rng.normal(size=8)
        f = lambda x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z + c[7] * x * y * z
        p = rng.uniform(0, sh - 1)
        idx, w = O.interp_coef(sh, p)
        assert abs(sum(w[n] * f(*idx[n]) for n in range(8)) - f(*p)) < 1e-10
        idx, dw = O.deriv_coef(sh, p)
        x, y, z = p
        grad = [c[1] + c[4] * y + c[6] * z + c[7] * y * z, c[2] + c[4] * x + c[5] * z + c[7] * x * z,
                c[3] + c[5] * y + c[6] * x + c[7] * x * y]
        for a in range(3):
            assert abs(sum(dw[a][n] * f(*idx[n]) for n in range(8)) - grad[a]) < 1e-10


def fields(w, h, d, fill=0.0, active=0):
    shapes = O.face_shapes(w, h, d)
    vel = [np.full(s, fill, np.float32) for s in shapes]
    act = [np.full(s, active, np.uint8) for s in shapes]
    return vel, act


def test_extrapolation_active_set_is_the_dilation():
    rng = np.random.default_rng(5)
    w, h, d = 9, 7, 11
    vel, act = fields(w, h, d)
    for a in range(3):
        act[a][...] = rng.uniform(size=act[a].shape) < 0.04
        vel[a][...] = rng.normal(size=vel[a].shape)
    before = [x.copy() for x in act]
    vb = [x.copy() for x in vel]
    for width in (1, 2, 3):
        va, aa = [x.copy() for x in vb], [x.copy() for x in before]
        n = O.extrapolate(w, h, d, va, aa, width)
        cross = ndimage.generate_binary_structure(3, 1)
        total = 0
        for a in range(3):
            want = ndimage.binary_dilation(before[a].astype(bool), structure=cross, iterations=width)
            assert np.array_equal(aa[a].astype(bool), want)
            total += int(want.sum() - before[a].sum())
            # the originally active faces keep their values
            assert np.array_equal(va[a][before[a] == 1], vb[a][before[a] == 1])
        assert n == total


def test_extrapolation_passes_read_only_the_previous_pass():
    w, h, d = 8, 1, 1
    vel, act = fields(w, h, d)
    u, au = vel[0], act[0]
    u[0, 0, 0], au[0, 0, 0] = 1.0, 1
    u[4, 0, 0], au[4, 0, 0] = 2.0, 1
    O.extrapolate(w, h, d, vel, act, 1)
    assert list(au[:, 0, 0]) == [1, 1, 0, 1, 1, 1, 0, 0, 0]
    assert u[1, 0, 0] == 1.0 and u[3, 0, 0] == 2.0 and u[5, 0, 0] == 2.0
    O.extrapolate(w, h, d, vel, act, 1)
    assert au[2, 0, 0] == 1 and u[2, 0, 0] == 1.5   # mean of its two neighbours from the last pass
    assert u[6, 0, 0] == 2.0


def test_extrapolation_copies_across_a_half_space():
    w, h, d = 10, 6, 5
    rng = np.random.default_rng(8)
    vel, act = fields(w, h, d)
    for a in range(3):
        prof = rng.normal(size=(1,) + vel[a].shape[1:]).astype(np.float32)
        vel[a][...] = prof
        act[a][:4] = 1
        vel[a][4:] = 123.0  # garbage on inactive faces is never read
    O.extrapolate(w, h, d, vel, act, 3)
    for a in range(3):
        assert act[a][:7].all() and not act[a][7:].any()
        assert np.array_equal(vel[a][:7], np.broadcast_to(vel[a][:1], vel[a][:7].shape))


def test_constant_field_extrapolates_to_itself():
    w, h, d = 6, 6, 6
    vel, act = fields(w, h, d, fill=0.5)
    for a in range(3):
        act[a][2:4, 2:4, 2:4] = 1
    O.extrapolate(w, h, d, vel, act, 4)
    for a in range(3):
        assert (vel[a][act[a] == 1] == 0.5).all()


def plane_solid(w, h, d, normal, offset):
    """Signed distance to the plane n.x = offset (world units, dx = 1/max), sampled at nodes."""
    dx = min(1.0 / w, 1.0 / h, 1.0 / d)
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    i, j, k = np.meshgrid(np.arange(w + 1), np.arange(h + 1), np.arange(d + 1), indexing="ij")
    return (n[0] * i * dx + n[1] * j * dx + n[2] * k * dx - offset).astype(np.float32), dx, n


def face_centres(w, h, d, a):
    s = O.face_shapes(w, h, d)[a]
    i, j, k = np.meshgrid(np.arange(s[0]), np.arange(s[1]), np.arange(s[2]), indexing="ij")
    return i + 0.5 * (a != 0), j + 0.5 * (a != 1), k + 0.5 * (a != 2)


@pytest.mark.parametrize("normal", [(0, 1, 0), (1, 1, 0), (1, -2, 3)])
def test_constraint_removes_the_incoming_normal_component(normal):
    w, h, d = 12, 12, 12
    solid, dx, n = plane_solid(w, h, d, normal, 0.45)
    u0 = np.array([0.3, -0.8, 0.25])
    if u0 @ n > 0:
        u0 = -u0
    vel, act = fields(w, h, d, active=1)
    for a in range(3):
        vel[a][...] = np.float32(u0[a])
    assert O.constrain_velocity(w, h, d, vel, act, solid) == 1
    un = u0 @ n
    for a in range(3):
        x, y, z = face_centres(w, h, d, a)
        phi = (n[0] * x + n[1] * y + n[2] * z) * dx - 0.45
        inside = phi < -1e-6
        outside = phi > 1e-6
        idx = [x, y, z][a]
        interior = (idx > 0.5) & (idx < [w, h, d][a] - 0.5)
        want = u0[a] - n[a] * un
        got = vel[a]
        assert np.allclose(got[inside & interior], want, atol=1e-5), a
        assert (got[outside & interior] == np.float32(u0[a])).all(), a


def test_constraint_keeps_a_leaving_velocity_and_applies_boundary_rule():
    w, h, d = 8, 8, 8
    solid, dx, n = plane_solid(w, h, d, (0, 1, 0), 0.3)
    vel, act = fields(w, h, d, active=1)
    for a, val in enumerate((-0.7, 0.9, 0.6)):
        vel[a][...] = val
    O.constrain_velocity(w, h, d, vel, act, solid)
    # v leaves the floor (u.n > 0): unchanged except where it leaves the domain at the top
    assert (vel[1][:, 1:h, :] == np.float32(0.9)).all() and (vel[1][:, h, :] == 0).all()
    assert (vel[0][0] == 0).all() and (vel[0][1:] == np.float32(-0.7)).all()   # outflow at x = 0
    assert (vel[2][:, :, d] == 0).all() and (vel[2][:, :, :d] == np.float32(0.6)).all()


def test_constraint_needs_a_solid_and_skips_inactive_faces():
    w, h, d = 6, 5, 7
    vel, act = fields(w, h, d, fill=-1.0, active=1)
    nosolid = np.full((w + 1, h + 1, d + 1), 0.2, np.float32)
    assert O.constrain_velocity(w, h, d, vel, act, nosolid) == 0
    assert all((v == -1.0).all() for v in vel)   # not even the boundary rule
    solid, dx, n = plane_solid(w, h, d, (0, 1, 0), 0.5)
    for a in range(3):
        act[a][...] = 0
        act[a][::2] = 1
    before = [v.copy() for v in vel]
    O.constrain_velocity(w, h, d, vel, act, solid)
    for a in range(3):
        assert np.array_equal(vel[a][act[a] == 0], before[a][act[a] == 0])
