"""Pins the CPU oracle (oracle/shkz_oracle.c) before it is trusted as the checker.

Each oracle function is checked against what fixes it independently of the oracle's own formula:
  * the reference's own header code (utility.h, sparse_matrix.h) -- live through oracle/_ref when
    the reference tree is present, and always through the committed fixtures that
    tests/golden/make_golden.py recorded from it;
  * closed forms (areas of half-planes and triangles, linear level-set crossings);
  * invariants of the assembled operator (symmetry, M-matrix, row sums = Dirichlet sums);
  * brute force (dense numpy solves on small grids);
  * the reference's known-answer test accuracytest3 (src/examples/accuracytest3-example.cpp:68-145):
    second-order convergence of the ghost-fluid pressure to |x-c|^2 - r^2.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
DP = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


# ------------------------------------------------------------------ geometry helpers
def test_get_area_matches_reference_fixtures(gold):
    q = np.ascontiguousarray(gold["quads"])
    got = np.array([O.lib().oracle_get_area(row.ctypes.data_as(DP)) for row in q])
    assert np.array_equal(got, gold["area"]), "get_area differs from utility::get_area bit for bit"


def test_fraction_matches_reference_fixtures(gold):
    got = np.array([O.lib().oracle_fraction(a, b) for a, b in gold["pairs"]])
    assert np.array_equal(got, gold["frac"])


def test_box_matches_reference_fixtures(gold):
    P, B0, B1 = (np.ascontiguousarray(gold[k]) for k in ("box_p", "box_p0", "box_p1"))
    got = np.array([O.lib().oracle_box_sdf(P[i].ctypes.data_as(DP), B0[i].ctypes.data_as(DP), B1[i].ctypes.data_as(DP))
                    for i in range(len(P))])
    assert np.array_equal(got, gold["box"])


@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="reference tree not present (oracle/_ref not built)")
def test_geometry_matches_live_reference():
    rng = np.random.default_rng(7)
    R = O.ref()
    for _ in range(3000):
        q = np.ascontiguousarray(rng.normal(size=4) * 10.0 ** rng.integers(-6, 2), np.float64)
        assert O.lib().oracle_get_area(q.ctypes.data_as(DP)) == R.ref_get_area(q.ctypes.data_as(DP))
        a, b = rng.normal(size=2)
        assert O.lib().oracle_fraction(a, b) == R.ref_fraction(a, b)


def area(q00, q01, q10, q11):
    q = np.array([q00, q01, q10, q11], np.float64)
    return O.lib().oracle_get_area(q.ctypes.data_as(DP))


def test_get_area_closed_forms():
    assert area(1, 1, 1, 1) == 0.0
    assert area(-1, -1, -1, -1) == 1.0
    # half plane x < 1/2: corners (0,0),(0,1) negative, (1,0),(1,1) positive
    assert area(-1, -1, 1, 1) == pytest.approx(0.5, abs=1e-15)
    # half plane crossing at x = 1/4
    assert area(-1, -1, 3, 3) == pytest.approx(0.25, abs=1e-15)
    # one negative corner: triangle with legs a, b -> ab/2
    assert area(-1, 1, 1, 3) == pytest.approx(0.5 * 0.5 * 0.5, abs=1e-15)
    # three negative corners = 1 - triangle
    assert area(-1, -1, -1, 1) == pytest.approx(1.0 - 0.125, abs=1e-15)
    # symmetry under relabelling x <-> y of a linear field
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, bx, by = rng.normal(size=3)
        f = lambda x, y: a + bx * x + by * y
        g = lambda x, y: a + by * x + bx * y
        assert area(f(0, 0), f(0, 1), f(1, 0), f(1, 1)) == pytest.approx(area(g(0, 0), g(0, 1), g(1, 0), g(1, 1)),
                                                                         abs=1e-12)


def test_get_area_linear_field_exact():
    """For a linear level set the marching polygon is exact: compare with a fine quadrature."""
    rng = np.random.default_rng(1)
    xs = (np.arange(400) + 0.5) / 400
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    for _ in range(20):
        a, bx, by = rng.normal(size=3)
        f = a + bx * X + by * Y
        exact = (f < 0).mean()
        v = lambda x, y: a + bx * x + by * y
        assert area(v(0, 0), v(0, 1), v(1, 0), v(1, 1)) == pytest.approx(exact, abs=5e-3)


def test_fraction_closed_forms():
    f = O.lib().oracle_fraction
    assert f(-1, -2) == 1.0 and f(1, 2) == 0.0 and f(0, 0) == 0.0 and f(-1, 0) == 1.0
    assert f(-1, 3) == pytest.approx(0.25)
    assert f(3, -1) == pytest.approx(0.25)  # symmetric: the liquid side length
    assert f(-0.5, 0.5) == 0.5


# ------------------------------------------------------------------ sparse matrix / SpMV
def test_csr_and_spmv_match_reference_sparse_matrix(gold):
    """The oracle's CSR rows (sorted columns) and SpMV equal Bridson's SparseMatrix +
    FixedSparseMatrix::multiply built by the reference code from the same triplets."""
    n = int(gold["sp_n"])
    rowptr = gold["sp_rowstart"]
    col, val = gold["sp_colindex"], gold["sp_value"]
    sysd = {"n": n, "rowptr": np.ascontiguousarray(rowptr), "col": np.ascontiguousarray(col),
            "val": np.ascontiguousarray(val)}
    assert np.array_equal(O.spmv(sysd, gold["sp_x"]), gold["sp_y"])
    # the same scene assembled by the oracle gives the same CSR
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 12, seed=5)
    mine = O.assemble(sc)
    assert mine["n"] == n
    assert np.array_equal(mine["rowptr"], rowptr)
    assert np.array_equal(mine["col"], col)
    assert np.array_equal(mine["val"], val)


# ------------------------------------------------------------------ assembly invariants
@pytest.mark.parametrize("kind,n,seed", [(S.SCENE_DROP, 24, 0), (S.SCENE_DROP_SOLIDS, 24, 1),
                                         (S.SCENE_SPHERES_BOX, 24, 2), (S.SCENE_RANDOM_DROP, 20, 3)])
def test_operator_invariants(kind, n, seed):
    sc = S.make_scene(kind, n, seed=seed)
    A = O.assemble(sc)
    N = A["n"]
    assert N > 0
    import scipy.sparse as sp
    M = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(N, N))
    assert abs(M - M.T).max() == 0.0                       # symmetric (same w computed from both sides)
    d = M.diagonal()
    assert (d > 0).all()
    off = M - sp.diags(d)
    assert off.max() <= 0.0                                 # M-matrix sign pattern
    rowsum = np.asarray(M.sum(axis=1)).ravel()
    assert (rowsum >= -1e-12 * d).all()                     # weak diagonal dominance
    # row sum = sum of conductances to air neighbours; check it cell by cell from the fractions
    area, rho = A["area"], A["rho"]
    im = A["index_map"]
    dx, dt = sc.dx, sc.dt
    idx = np.argwhere(im >= 0)
    rng = np.random.default_rng(seed)
    for (i, j, k) in idx[rng.choice(len(idx), size=min(200, len(idx)), replace=False)]:
        s = 0.0
        for dim, (di, dj, dk) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
            for sgn in (1, -1):
                q = (i + sgn * di, j + sgn * dj, k + sgn * dk)
                if min(q) < 0 or q[0] >= sc.w or q[1] >= sc.h or q[2] >= sc.d:
                    continue
                f = q if sgn > 0 else (i, j, k)
                a, r = float(area[dim][f]), float(rho[dim][f])
                if a and r and not (sc.fluid[q] < 0):
                    s += dt * a / (dx * dx * r)
        assert rowsum[im[i, j, k]] == pytest.approx(s, rel=1e-12, abs=1e-9)


def test_unknown_labelling_rules():
    """Isolated liquid cells and liquid cells cut off by solid are not unknowns."""
    w = h = d = 6
    fluid = np.full((w, h, d), 1.0)
    fluid[1, 1, 1] = -0.5                # isolated single cell -> excluded
    fluid[3:5, 3:5, 3:5] = -0.5          # 2x2x2 block -> 8 unknowns
    sc = S.Scene(w, h, d, 1.0 / 120, fluid, None, *(np.zeros(s, np.float32) for s in O.face_shapes(w, h, d)))
    A = O.assemble(sc)
    assert A["n"] == 8 and A["index_map"][1, 1, 1] == -1


# ------------------------------------------------------------------ solver
def test_cg_matches_dense_solve():
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 14, seed=11)
    A = O.assemble(sc)
    import scipy.sparse as sp
    N = A["n"]
    M = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(N, N)).toarray()
    x_dense = np.linalg.solve(M, A["b"])
    for mode in (O.MIC_NONE, O.MIC_FAITHFUL, O.MIC_APPLIED):
        x, it, rr, ok = O.pcg(A, residual=1e-13, mic_mode=mode)
        assert ok
        assert np.abs(x - x_dense).max() <= 1e-9 * np.abs(x_dense).max()
        assert np.abs(A["b"] - M @ x).max() <= 1e-12 * np.abs(A["b"]).max() * 10


def test_faithful_mic_has_plain_cg_iterates():
    """The reference discards MIC(0) (pcg_solver.h:383): iterates equal plain CG exactly."""
    sc = S.make_scene(S.SCENE_DROP, 16, seed=0)
    A = O.assemble(sc)
    x0, it0, rr0, _ = O.pcg(A, residual=1e-9, mic_mode=O.MIC_NONE)
    x1, it1, rr1, _ = O.pcg(A, residual=1e-9, mic_mode=O.MIC_FAITHFUL)
    assert it0 == it1 and rr0 == rr1 and np.array_equal(x0, x1)


def test_mic0_is_exact_cholesky_without_fill():
    """On a tridiagonal SPD matrix incomplete Cholesky has no dropped fill, so MIC(0) is the
    exact factor and PCG converges in one iteration."""
    n = 50
    rowptr, col, val = [0], [], []
    for i in range(n):
        if i > 0:
            col.append(i - 1); val.append(-1.0)
        col.append(i); val.append(2.5 + 0.01 * i)
        if i < n - 1:
            col.append(i + 1); val.append(-1.0)
        rowptr.append(len(col))
    sysd = {"n": n, "rowptr": np.array(rowptr, np.int64), "col": np.array(col, np.int64),
            "val": np.array(val, np.float64), "b": np.linspace(1, 2, n)}
    x, it, rr, ok = O.pcg(sysd, residual=1e-12, mic_mode=O.MIC_APPLIED)
    assert ok and it == 1


def test_empty_system_converges_immediately():
    w = h = d = 4
    fluid = np.full((w, h, d), 1.0)
    sc = S.Scene(w, h, d, 1.0 / 120, fluid, None, *(np.ones(s, np.float32) for s in O.face_shapes(w, h, d)))
    r, P, P64 = O.project(sc, O.default_params(residual=1e-8))
    assert r.n_unknowns == 0 and r.iterations == 0 and not P.any()
    # no liquid: interior faces keep their velocity (P = 0), boundary faces are zeroed
    assert (sc.u[1:-1] == 1).all() and (sc.u[0] == 0).all() and (sc.u[-1] == 0).all()


# ------------------------------------------------------------------ velocity update
def test_velocity_update_removes_divergence():
    """After the update, div(u) on the unknowns equals -(b - A p) up to float rounding."""
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 20, seed=4)
    A = O.assemble(sc)
    x, it, rr, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_APPLIED)
    P = O.scatter(A, x, (sc.w, sc.h, sc.d)).astype(np.float32)
    before, _ = O.divergence(sc, A["index_map"], A["area"])
    O.velocity_update(sc, P, A["area"], A["rho"])
    after, div = O.divergence(sc, A["index_map"], A["area"])
    assert before > 1.0
    assert after < 1e-4 * before
    # exact identity in FP64 with the float pressure
    Pd = O.gather(A, P.astype(np.float64))
    resid = A["b"] - O.spmv(A, Pd)
    d = O.gather(A, div)
    assert np.abs(d + resid).max() <= 2e-5 * max(1.0, np.abs(A["b"]).max())


# ------------------------------------------------------------------ the reference KAT
def accuracy_scene(n, r):
    """accuracytest3: u = 2(x - c) on faces, phi = |x - c| - r (float), dt = 1, no solid."""
    dx = 1.0 / n
    c = 0.5
    ii = (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(ii, ii, ii, indexing="ij")
    fluid = (np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2) - r).astype(np.float32)
    fi = np.arange(n + 1) * dx
    u = np.broadcast_to(2.0 * (fi - c)[:, None, None], (n + 1, n, n)).astype(np.float32).copy()
    v = np.broadcast_to(2.0 * (fi - c)[None, :, None], (n, n + 1, n)).astype(np.float32).copy()
    w = np.broadcast_to(2.0 * (fi - c)[None, None, :], (n, n, n + 1)).astype(np.float32).copy()
    return S.Scene(n, n, n, 1.0, np.ascontiguousarray(fluid), None, u, v, w), (X, Y, Z)


def accuracy_error(n, r0=0.4, trials=4):
    dx = 1.0 / n
    worst = 0.0
    qs = [q for q in range(-trials, trials + 1) if q] + [0]
    for q in qs:
        r = r0 + math.sqrt(3.0) * dx / trials * q
        sc, (X, Y, Z) = accuracy_scene(n, r)
        prm = O.default_params(residual=1e-12, eps_fluid=1e-18, max_iterations=20000)
        res, P, _ = O.project(sc, prm, pressure64=False)
        exact = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 - r * r
        mask = P != 0
        worst = max(worst, np.abs(exact[mask] - P[mask]).max())
    return worst


def test_accuracy_known_answer_second_order():
    """The reference's accuracytest3: the ghost-fluid projection recovers p = |x-c|^2 - r^2 with
    second-order accuracy (the reference logs order = log2(prev/curr), expected about 2)."""
    e1 = accuracy_error(16)
    e2 = accuracy_error(32)
    order = math.log2(e1 / e2)
    assert e2 < 2e-3
    assert order > 1.6, (e1, e2, order)


def test_permuted_system_is_the_same_system():
    """O.permuted renumbers the unknowns only: the permuted system's solution is the permuted
    solution, and its matrix-vector product is the permuted product (so the spread of CG counts
    over orders measures rounding alone)."""
    A = O.assemble(S.make_scene(S.SCENE_DROP_SOLIDS, 16, seed=1))
    n = A["n"]
    rng = np.random.default_rng(3)
    perm = rng.permutation(n)
    Ap = O.permuted(A, perm)
    v = rng.standard_normal(n)
    assert np.allclose(O.spmv(Ap, v[perm]), O.spmv(A, v)[perm], rtol=0, atol=1e-12)
    x, _, _, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_NONE)
    xp, _, _, okp = O.pcg(Ap, residual=1e-12, mic_mode=O.MIC_NONE)
    assert ok and okp
    assert np.linalg.norm(xp - x[perm]) <= 1e-9 * np.linalg.norm(x)
    lo, hi = O.cg_count_envelope(A, residual=1e-8, n_orders=2)
    _, it, _, _ = O.pcg(A, residual=1e-8, mic_mode=O.MIC_NONE)
    assert lo < it < hi
