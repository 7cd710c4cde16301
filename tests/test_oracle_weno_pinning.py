"""Pins the WENO option of the advection oracle (oracle/shkz_oracle_adv.c: oracle_weno6,
oracle_weno3 and the weno flag of the advection passes; the reference's WENO::interp6,
include/shiokaze/math/WENO.h:55-88, WENO3::interpolate, include/shiokaze/math/WENO3.h:105-135,
and macadvection3.cpp:85-88,200-203).

  * interp6 equals the reference's own WENO::interp6, compiled from its header into oracle/_ref,
    bit for bit on random, rounded, steep and discontinuous samples (where that library is built);
  * closed forms that hold whatever the build: a constant is reproduced; every cubic is reproduced
    (each candidate stencil is exact on cubics); on samples so small that beta^2 << eps the
    weights fall back to the linear ones and the blend is the quintic through all six nodes (a
    wrong linear weight fails it); next to a jump the stencils that miss it carry the result (a
    smoothness indicator reading the wrong samples or a wrong sign fails it); on smooth data the
    error falls at least like h^4;
  * the 3D sampling reduces x, then y, then z over the clamped 6x6x6 block, with positions clamped
    into the array: checked against the composition of 1D calls (to the reference's own interp6
    where built), at nodes it returns the sample, and data that vary along one axis only give the
    1D result;
  * the advection passes with the flag: zero velocity changes nothing, a uniform field stays
    uniform, a linear scalar moves by u dt / dx, MacCormack stays within its trilinear bounds and
    keeps the semi-Lagrangian value in the narrow band, and WENO differs from trilinear on rough
    data.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

EPS = np.finfo(np.float64).eps


def bits(x):
    return np.float64(x).view(np.uint64)


def sample_sets(n, seed):
    rng = np.random.default_rng(seed)
    for t in range(n):
        f = rng.normal(size=6) * 10.0 ** rng.uniform(-4, 4)
        if t % 3 == 0:
            f = np.round(f)
        if t % 7 == 0:
            f[rng.integers(0, 6):] += rng.normal() * 1e3
        if t % 11 == 0:
            f = np.exp(rng.uniform(0, 5) * np.arange(-2, 4))
        x = float(rng.uniform()) if t % 5 else float(rng.integers(0, 2))
        yield x, f


@pytest.mark.skipif(O.ref() is None, reason="reference driver not built (no reference tree)")
def test_weno6_matches_reference_bitwise():
    bad = [(x, f) for x, f in sample_sets(20000, 1) if bits(O.weno6(x, f)) != bits(O.ref_weno6(x, f))]
    assert not bad, bad[:3]


def test_weno6_constant():
    for c in (0.0, 1.0, -3.25, 1e-7, 12345.678):
        for x in (0.0, 0.3, 0.5, 1.0):
            r = O.weno6(x, np.full(6, c))
            assert abs(r - c) <= 4 * EPS * abs(c), (c, x, r)


def test_weno6_reproduces_cubics():
    rng = np.random.default_rng(2)
    t = np.arange(-2, 4, dtype=np.float64)
    for _ in range(200):
        a = rng.normal(size=4)
        f = a[0] + a[1] * t + a[2] * t ** 2 + a[3] * t ** 3
        x = float(rng.uniform())
        want = a[0] + a[1] * x + a[2] * x ** 2 + a[3] * x ** 3
        assert abs(O.weno6(x, f) - want) <= 1e-12 * (1 + np.abs(f).max()), (a, x)


def test_weno6_linear_weights_give_the_quintic():
    """beta ~ 1e-26, so beta^2 << eps: the weights are the linear ones, gamma_r / sum gamma."""
    t = np.arange(-2, 4, dtype=np.float64)
    rng = np.random.default_rng(3)
    for _ in range(50):
        a = rng.normal(size=6)
        f = 1e-13 * sum(a[n] * t ** n for n in range(6))
        x = float(rng.uniform())
        want = 1e-13 * sum(a[n] * x ** n for n in range(6))
        assert abs(O.weno6(x, f) - want) <= 1e-10 * 1e-13 * np.abs(a).sum(), (a, x)
        # a blend of the cubics alone (any other weights) misses the quintic term
        cub = 1e-13 * sum(a[n] * x ** n for n in range(4))
        assert abs(want - cub) > 1e-3 * 1e-13 * abs(a[4] * x ** 4 + a[5] * x ** 5) or abs(a[4]) < 1e-3


def test_weno6_jump_takes_the_smooth_stencils():
    for x in (0.1, 0.5, 0.9):
        # jump between nodes 2 and 3: stencils 0 and 1 (nodes -2..2) see a constant
        assert abs(O.weno6(x, np.array([0, 0, 0, 0, 0, 1.0]))) < 1e-14
        # jump between nodes -2 and -1: stencils 1 and 2 (nodes -1..3) see a constant
        assert abs(O.weno6(x, np.array([1.0, 0, 0, 0, 0, 0]))) < 1e-14
        # a jump inside the cell [0, 1] stays within the data and monotone samples stay monotone
        r = O.weno6(x, np.array([0, 0, 0, 1.0, 1.0, 1.0]))
        assert -1e-3 < r < 1 + 1e-3


def test_weno6_order_on_smooth_data():
    errs = []
    for h in (0.2, 0.1, 0.05, 0.025):
        e = 0.0
        for x0 in np.linspace(0.1, 2.0, 7):
            f = np.sin(x0 + h * np.arange(-2, 4))
            for x in (0.25, 0.5, 0.75):
                e = max(e, abs(O.weno6(x, f) - math.sin(x0 + h * x)))
        errs.append(e)
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(rates) > 3.8, (errs, rates)


def compose(U, p, f1):
    """WENO3::interpolate(array, p), order 6, from 1D calls f1 (WENO3.h:105-135)."""
    s = U.shape
    c = [max(0.0, min(s[a] - 1.0, p[a])) for a in range(3)]
    b = [int(min(c[a], s[a] - 2.0)) for a in range(3)]
    fr = [c[a] - b[a] for a in range(3)]
    cl = [lambda v, n=s[a]: min(max(v, 0), n - 1) for a in range(3)]
    vz = []
    for kk in range(6):
        vy = []
        for jj in range(6):
            vx = [float(U[cl[0](b[0] + ii - 2), cl[1](b[1] + jj - 2), cl[2](b[2] + kk - 2)]) for ii in range(6)]
            vy.append(f1(fr[0], np.array(vx)))
        vz.append(f1(fr[1], np.array(vy)))
    return f1(fr[2], np.array(vz))


def test_weno3_composition_and_clamping():
    rng = np.random.default_rng(5)
    U = rng.normal(size=(7, 5, 9)).astype(np.float32)
    U[3:, :, 4:] += 4.0  # a jump
    f1 = O.ref_weno6 if O.ref() is not None else O.weno6
    pts = [rng.uniform(-2, 11, size=3) for _ in range(300)]
    pts += [np.array([6.0, 4.0, 8.0]), np.array([0.0, 0.0, 0.0]), np.array([5.5, 3.999, 7.25])]
    for p in pts:
        assert bits(O.weno3(U, p)) == bits(compose(U, p, f1)), p


def test_weno3_nodes_clamping_and_separable_data():
    rng = np.random.default_rng(6)
    U = rng.normal(size=(6, 7, 5)).astype(np.float32)
    for i, j, k in [(0, 0, 0), (2, 3, 1), (5, 6, 4), (4, 1, 3)]:
        assert abs(O.weno3(U, [i, j, k]) - U[i, j, k]) <= 1e-6 * (1 + abs(U[i, j, k]))
    # outside the array: the position is clamped onto it
    assert O.weno3(U, [-3.0, 2.5, 1.25]) == O.weno3(U, [0.0, 2.5, 1.25])
    assert O.weno3(U, [9.0, 10.0, 1.25]) == O.weno3(U, [5.0, 6.0, 1.25])
    # data along x only: the y and z reductions see constants
    g = rng.normal(size=8).astype(np.float32)
    V = np.broadcast_to(g[:, None, None], (8, 6, 6)).copy()
    for x in (0.3, 2.7, 4.5, 6.9):
        base = min(int(x), 6)
        f = np.array([g[min(max(base + n - 2, 0), 7)] for n in range(6)], np.float64)
        want = O.weno6(x - base, f)
        assert abs(O.weno3(V, [x, 2.2, 3.7]) - want) <= 8 * EPS * (1 + abs(want))


# ----------------------------------------------------------------------------- advection passes
def faces(w, h, d, fill=0.0, active=1):
    shapes = O.face_shapes(w, h, d)
    return [np.full(s, fill, np.float32) for s in shapes], [np.full(s, active, np.uint8) for s in shapes]


def deep_fluid(w, h, d):
    return np.full((w, h, d), -1.0, np.float32)


@pytest.mark.parametrize("mc", [False, True])
def test_weno_advection_zero_and_uniform(mc):
    w, h, d = 9, 8, 7
    rng = np.random.default_rng(1)
    vel, act = faces(w, h, d)
    q = rng.normal(size=(w, h, d)).astype(np.float32)
    q0 = q.copy()
    O.advect_scalar(w, h, d, vel, act, q, 0.1, fluid=deep_fluid(w, h, d), maccormack=mc, weno=True)
    assert np.array_equal(q, q0)
    w = h = d = 10
    vel, act = faces(w, h, d)
    for a, c in enumerate((0.5, -0.25, 0.75)):
        vel[a][...] = c
    O.advect_velocity(w, h, d, vel, act, deep_fluid(w, h, d), 0.05, dx=0.1, maccormack=mc, weno=True)
    for a, c in enumerate((0.5, -0.25, 0.75)):
        assert np.abs(vel[a] - np.float32(c)).max() <= 2 * abs(np.spacing(np.float32(c))), a


@pytest.mark.parametrize("mc", [False, True])
def test_weno_linear_scalar_moves(mc):
    w, h, d = 20, 18, 8
    dx = 1.0 / 16
    vel, act = faces(w, h, d)
    vel[0][...] = 0.5
    vel[1][...] = -0.25
    dt = 0.0625 * 0.7  # shift (0.35, -0.175, 0) cells
    i, j, k = np.meshgrid(np.arange(w), np.arange(h), np.arange(d), indexing="ij")
    q = (2.0 * i + 4.0 * j + 1.0 * k).astype(np.float32)
    O.advect_scalar(w, h, d, vel, act, q, dt, fluid=deep_fluid(w, h, d), dx=dx, maccormack=mc, weno=True)
    sx, sy = np.float32(np.float32(0.5 * dt) / dx), np.float32(np.float32(-0.25 * dt) / dx)
    want = 2.0 * (i - float(sx)) + 4.0 * (j - float(sy)) + k
    # the clamped stencil reaches 3 cells in, twice over for MacCormack's backward pass
    m = 7 if mc else 4
    inner = (i >= m) & (i <= w - 1 - m) & (j >= m) & (j <= h - 1 - m)
    assert np.abs(q[inner] - want[inner]).max() < 1e-5


def test_weno_maccormack_bounds_narrow_band_and_difference():
    w = h = d = 14
    dx = 1.0 / 14
    rng = np.random.default_rng(4)
    vel, act = faces(w, h, d)
    for a in range(3):
        vel[a][...] = rng.uniform(-1, 1, size=vel[a].shape)
    q = (rng.uniform(size=(w, h, d)) > 0.5).astype(np.float32)
    fluid = np.where(np.arange(w)[:, None, None] < 7, -1.0, 0.5).astype(np.float32) * np.ones((w, h, d), np.float32)
    dt = 0.5 * dx
    q_sl, q_mc, q_tri = q.copy(), q.copy(), q.copy()
    O.advect_scalar(w, h, d, vel, act, q_sl, dt, fluid=fluid, dx=dx, maccormack=False, weno=True)
    O.advect_scalar(w, h, d, vel, act, q_mc, dt, fluid=fluid, dx=dx, maccormack=True, weno=True)
    O.advect_scalar(w, h, d, vel, act, q_tri, dt, fluid=fluid, dx=dx, maccormack=False, weno=False)
    # outside the narrow band (x <= 5: departure points stay in fluid = -1) MacCormack is clamped
    # to the trilinear stencil's min/max; inside it keeps the WENO value, which may overshoot
    assert (q_mc[:6] >= 0).all() and (q_mc[:6] <= 1).all()
    assert np.array_equal(q_mc[9:], q_sl[9:])
    assert not np.array_equal(q_mc[1:6], q_sl[1:6])
    assert not np.array_equal(q_sl, q_tri)
    # WENO on 0/1 data overshoots far less than the cubics it blends would
    assert q_sl.min() > -0.2 and q_sl.max() < 1.2
