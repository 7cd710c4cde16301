"""Pins the oracle's surface tension, PI volume controller and warm start
(src/projection/macpressuresolver3.cpp:85-115, :203-217 with :290, :221-242) by what fixes them
independently of the oracle's own formulas:

  * surface tension: second differences of quadratics and cubics are exact, so the 7-point
    curvature of phi = |x-c|^2 - R^2 is 6 everywhere and that of phi = (x-a)^3 - e is 6(x-a); the
    curvature a face sees is then the curvature at the interface, where the linear interpolant of
    phi crosses zero; a planar phi has none.  Physics: a static drop obeys Laplace's law, the
    pressure inside is 2 kappa / r, converging at second order, with vanishing parasitic velocity;
  * volume controller: zero when Gain or the target is 0, zero at the target with a fresh
    integrator, the integrator is the running sum of x dt, below the target the correction is
    positive, one value derived by hand; and through a whole projection the post-projection
    divergence of every unknown equals the correction (div' = div + (A p) = div + b = correction,
    since b = -div + correction);
  * warm start: the result is p_prev + (the solve of A y = b - A p_prev) and meets the tolerance
    relative to b - A p_prev; a warm start from the converged answer leaves it in place.
"""
import ctypes
import math

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O


def zero_vel(n):
    return [np.zeros(s, np.float32) for s in O.face_shapes(n, n, n)]


def cell_centres(n):
    ii = (np.arange(n) + 0.5) / n
    return np.meshgrid(ii, ii, ii, indexing="ij")


def apply_surface_tension(sc, kappa, eps_fluid=1e-18):
    """Runs oracle_surface_tension on sc's velocities (in place) with the oracle's fractions."""
    prm = O.default_params(eps_fluid=eps_fluid)
    area, rho = O.fractions(sc, prm)
    g, keep = O.make_grid(sc)
    R = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in rho])
    O.lib().oracle_surface_tension(ctypes.byref(g), kappa, R)
    return rho


def interior_faces(n, dim):
    """Faces whose two cells both have unclamped 7-point stencils."""
    shape = O.face_shapes(n, n, n)[dim]
    idx = np.indices(shape)
    ok = np.ones(shape, bool)
    for a in range(3):
        lo = 2 if a == dim else 1
        ok &= (idx[a] >= lo) & (idx[a] <= n - 2)
    return ok


def test_surface_tension_quadratic_levelset_has_curvature_six():
    n, kappa, dt = 24, 0.7, 1.0 / 120.0
    dx = 1.0 / n
    X, Y, Z = cell_centres(n)
    phi = ((X - 0.5) ** 2 + (Y - 0.45) ** 2 + (Z - 0.55) ** 2 - 0.3 ** 2).astype(np.float32)
    sc = S.Scene(n, n, n, dt, np.ascontiguousarray(phi), None, *zero_vel(n))
    rho = apply_surface_tension(sc, kappa)
    vel = [sc.u, sc.v, sc.wv]
    crossed = 0
    for dim in range(3):
        r = rho[dim]
        cut = (r > 0) & (r < 1)
        # faces not cut by the interface get no force
        assert (vel[dim][~cut] == 0).all()
        m = cut & interior_faces(n, dim)
        crossed += int(m.sum())
        # phi(i,j,k) < 0 on the face's high cell: the force points towards -e_dim
        idx = np.nonzero(m)
        hi_liquid = phi[idx] < 0
        sgn = np.where(hi_liquid, -1.0, 1.0)
        # the curvature the face sees, recovered from the force: it must be 6 (exact for a quadratic)
        c = -vel[dim][idx].astype(np.float64) * dx * r[idx] / (dt * kappa) * sgn
        assert np.abs(c - 6.0).max() < 6.0 * 1e-3, np.abs(c - 6.0).max()
    assert crossed > 500


def test_surface_tension_cubic_levelset_curvature_at_the_interface():
    """phi = (x-a)^3 - e: curvature 6(x-a) at cell centres (exact); a face sees it interpolated to
    the interface, i.e. 6(x_G - a) with x_G the zero of the linear interpolant of phi between the
    two cell centres -- whichever side holds the liquid."""
    n, kappa, dt, a, e = 32, 1.3, 1.0 / 120.0, 0.5, 0.1 ** 3
    dx = 1.0 / n
    X, Y, Z = cell_centres(n)
    phi = ((X - a) ** 3 - e).astype(np.float32)
    sc = S.Scene(n, n, n, dt, np.ascontiguousarray(phi), None, *zero_vel(n))
    # the interface x = a + 0.1 runs between x-faces; flip the liquid side in a second scene
    for flip in (1.0, -1.0):
        f = np.ascontiguousarray(flip * phi)
        sc = S.Scene(n, n, n, dt, f, None, *zero_vel(n))
        rho = apply_surface_tension(sc, kappa)
        assert (sc.v == 0).all() and (sc.wv == 0).all()
        r = rho[0]
        idx = np.nonzero((r > 0) & (r < 1))
        assert len(idx[0]) == n * n  # one x-face per (j, k) row
        i = idx[0]
        x_hi, x_lo = (i + 0.5) * dx, (i - 0.5) * dx
        p_hi, p_lo = f[idx].astype(np.float64), f[(i - 1, idx[1], idx[2])].astype(np.float64)
        x_g = x_hi + (x_lo - x_hi) * p_hi / (p_hi - p_lo)
        expect = flip * 6.0 * (x_g - a)  # the Laplacian of flip*phi
        sgn = np.where(p_hi < 0, -1.0, 1.0)
        c = -sc.u[idx].astype(np.float64) * dx * r[idx] / (dt * kappa) * sgn
        assert np.abs(c - expect).max() < 1e-3 * np.abs(expect).max(), np.abs(c - expect).max()


def test_surface_tension_plane_has_no_force():
    n, kappa, dt = 20, 2.0, 1.0 / 120.0
    X, Y, Z = cell_centres(n)
    phi = (0.3 * X - 0.8 * Y + 0.5 * Z - 0.01).astype(np.float32)
    sc = S.Scene(n, n, n, dt, np.ascontiguousarray(phi), None, *zero_vel(n))
    rho = apply_surface_tension(sc, kappa)
    dx = 1.0 / n
    for dim, u in enumerate((sc.u, sc.v, sc.wv)):
        m = interior_faces(n, dim) & (rho[dim] > 0) & (rho[dim] < 1)
        assert m.sum() > 50
        # float rounding of the second differences only: far below one unit of curvature
        scale = dt * kappa / (dx * rho[dim][m])
        assert (np.abs(u[m]) <= 1e-3 * scale).all()


def laplace_drop(n, kappa=1.0, r=0.25):
    dx = 1.0 / n
    X, Y, Z = cell_centres(n)
    phi = (np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) - r).astype(np.float32)
    sc = S.Scene(n, n, n, 1.0 / 120.0, np.ascontiguousarray(phi), None, *zero_vel(n))
    prm = O.default_params(residual=1e-10, mic_mode=O.MIC_APPLIED, surface_tension=kappa)
    res, P, P64 = O.project(sc, prm)
    assert res.converged
    inside = phi < -3 * dx
    p_in = P64[inside]
    vmax = max(float(np.abs(a).max()) for a in (sc.u, sc.v, sc.wv))
    force_scale = sc.dt * kappa / (dx * r)
    return p_in.mean() / (2.0 * kappa / r) - 1.0, (p_in.max() - p_in.min()) / p_in.mean(), vmax / force_scale


def test_surface_tension_static_drop_obeys_laplace_law():
    e1, s1, v1 = laplace_drop(24)
    e2, s2, v2 = laplace_drop(48)
    assert abs(e1) < 5e-3 and abs(e2) < 1.5e-3, (e1, e2)
    assert math.log2(abs(e1) / abs(e2)) > 1.5, (e1, e2)
    assert s1 < 2e-3 and s2 < 2e-3
    # parasitic currents: a small fraction of one step's force, shrinking under refinement
    assert v1 < 2e-2 and v2 < v1, (v1, v2)


# ------------------------------------------------------------------ volume controller
def volume(cur, tgt, dt, gain, y):
    yy = ctypes.c_double(y)
    c = O.lib().oracle_volume_correction(cur, tgt, dt, gain, ctypes.byref(yy))
    return c, yy.value


def test_volume_correction_off_and_at_target():
    assert volume(0.9, 1.0, 0.1, 0.0, 0.0) == (0.0, 0.0)
    assert volume(0.9, 0.0, 0.1, 1.0, 0.0) == (0.0, 0.0)
    c, y = volume(1.0, 1.0, 0.1, 1.0, 0.0)
    assert c == 0.0 and y == 0.0


def test_volume_correction_integrator_and_sign():
    y = 0.0
    total = 0.0
    for cur, dt in ((0.9, 0.1), (0.95, 0.05), (1.02, 0.2)):
        c, y = volume(cur, 1.0, dt, 1.0, y)
        total += (cur - 1.0) / 1.0 * dt
        assert y == pytest.approx(total, rel=1e-15, abs=1e-18)
    # a fresh integrator below the target pushes the volume up, above the target down
    assert volume(0.8, 1.0, 0.1, 1.0, 0.0)[0] > 0
    assert volume(1.2, 1.0, 0.1, 1.0, 0.0)[0] < 0
    # reset (post_initialize sets the integrator to 0): the same call gives the same value again
    assert volume(0.9, 1.0, 0.1, 1.0, 0.0) == volume(0.9, 1.0, 0.1, 1.0, 0.0)


def test_volume_correction_hand_derived_value():
    """Gain 1, volume 0.9 of a target 1, dt 0.1, fresh integrator:
    x = (0.9 - 1)/1 = -0.1;  y = 0 + x dt = -0.01;  kp = 1 * 2.3 / (25 * 0.01) = 9.2;
    ki = kp^2 / 16 = 84.64 / 16 = 5.29;
    correction = -(kp x + ki y) / (x + 1) = -(-0.92 - 0.0529) / 0.9 = 0.9729 / 0.9 = 1.081."""
    c, y = volume(0.9, 1.0, 0.1, 1.0, 0.0)
    assert y == pytest.approx(-0.01, rel=1e-12)
    assert c == pytest.approx(1.081, rel=1e-12)


def test_volume_correction_sets_the_post_projection_divergence():
    """b = -div + correction on every row, so after the update div' = correction on every unknown."""
    sc = S.make_scene(S.SCENE_DROP, 24, seed=0)
    corr = 0.37
    prm = O.default_params(residual=1e-12, mic_mode=O.MIC_APPLIED, rhs_correct=corr)
    work = sc.copy()
    res, P, P64 = O.project(work, prm)
    assert res.converged
    A = O.assemble(sc)
    _, div = O.divergence(work, A["index_map"], A["area"])
    d = div[A["index_map"] >= 0]
    # float velocity storage: ~1e-7 of the pre-projection divergence
    tol = 2e-6 * np.abs(A["b"]).max()
    assert np.abs(d - corr).max() <= tol, np.abs(d - corr).max()


# ------------------------------------------------------------------ warm start
def test_warm_start_solves_for_the_difference():
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 20, seed=1)
    A = O.assemble(sc)
    n = A["n"]
    prev_cells = np.random.default_rng(3).normal(size=(20, 20, 20))
    prev_cells[A["index_map"] < 0] = 0.0
    p_prev = O.gather(A, prev_cells)
    for mode in (O.MIC_NONE, O.MIC_APPLIED):
        prm = O.default_params(residual=1e-9, mic_mode=mode)
        prm.warm_start = prev_cells.ctypes.data
        res, P, P64 = O.project(sc.copy(), prm)
        x = O.gather(A, P64)
        b_eff = A["b"] - O.spmv(A, p_prev)
        # the reference's semantics: x = p_prev + y with A y = b - A p_prev to the tolerance
        assert np.abs(A["b"] - O.spmv(A, x)).max() <= 1e-9 * np.abs(b_eff).max() * (1 + 1e-9)
        y, it, rr, ok = O.pcg(A, b=b_eff, residual=1e-9, mic_mode=mode)
        assert ok and it == res.iterations
        assert np.allclose(x, p_prev + y, rtol=0, atol=1e-12 * np.abs(x).max())
        assert n > 100


def test_warm_start_from_the_answer_stays_there():
    sc = S.make_scene(S.SCENE_DROP, 20, seed=2)
    A = O.assemble(sc)
    x_star, it, rr, ok = O.pcg(A, residual=1e-13, mic_mode=O.MIC_APPLIED)
    cells = O.scatter(A, x_star, (20, 20, 20))
    prm = O.default_params(residual=1e-6, mic_mode=O.MIC_APPLIED)
    prm.warm_start = cells.ctypes.data
    res, P, P64 = O.project(sc.copy(), prm)
    x = O.gather(A, P64)
    assert res.converged
    assert np.abs(x - x_star).max() <= 1e-9 * np.abs(x_star).max()
    # the tolerance is relative to the warm-started residual (already ~1e-13 of b), so the solve
    # still iterates, and it may only reduce that residual further
    r_star = np.abs(A["b"] - O.spmv(A, x_star)).max()
    assert np.abs(A["b"] - O.spmv(A, x)).max() <= 1e-6 * r_star * (1 + 1e-9) + 1e-15 * np.abs(A["b"]).max()
