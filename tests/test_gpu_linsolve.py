"""The LinSolver drop-in (shkz_linsolve_csr) against the oracle's pcg (the reference iteration) and
dense solves: the reference's pressure systems (3D and 2D), a generic irregular SPD matrix, and
the edge cases of pcg_solver.h:246-295 (zero rhs, singular pure-Neumann system, NaN, bad input)."""
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
def ls():
    s = S.LinSolver(0)
    yield s
    s.close()


def systems():
    yield "drop3d_32", O.assemble(S.make_scene(S.SCENE_DROP, 32, seed=0))
    yield "solids3d_28", O.assemble(S.make_scene(S.SCENE_DROP_SOLIDS, 28, seed=1))
    yield "box2d_96", O.assemble2(S.make_scene2d(S.SCENE_SPHERES_BOX, 96, seed=2))


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_reference_iteration_on_pressure_systems(ls, idx):
    name, A = list(systems())[idx]
    say(f"linsolve cg {name}")
    x_o, it_o, rr_o, ok = O.pcg(A, residual=1e-8, mic_mode=O.MIC_FAITHFUL)
    x, r = ls.solve(A["rowptr"], A["col"], A["val"], A["b"], residual=1e-8)
    assert r.status == 0 and r.converged and ok
    # CG iterates depend on dot-product summation order (blocked here, serial in the oracle)
    lo, hi = O.cg_count_envelope(A, residual=1e-8)  # the oracle's counts under other unknown orders
    assert lo <= r.iterations <= hi, (r.iterations, it_o, lo, hi)
    resid = A["b"] - O.spmv(A, x)
    assert np.abs(resid).max() <= 1e-8 * np.abs(A["b"]).max() * 1.0001
    assert r.reresid == pytest.approx(np.abs(resid).max() / np.abs(A["b"]).max(), rel=1e-6)
    assert np.linalg.norm(x - x_o) <= 1e-5 * np.linalg.norm(x_o)


def test_jacobi_pcg_matches_dense_solve_on_irregular_spd(ls):
    """A weighted graph Laplacian plus a positive diagonal, with rows of 1 to ~40 entries and
    unsorted columns: not a grid, not Poisson-shaped."""
    say("linsolve jacobi irregular")
    rng = np.random.default_rng(5)
    n = 3000
    rows, cols, vals = [], [], []
    for i in range(n):
        deg = 1 if i % 97 == 0 else rng.integers(1, 8)
        if i % 500 == 0:
            deg = 40
        for j in rng.choice(n, size=deg, replace=False):
            if j == i:
                continue
            wgt = rng.uniform(0.1, 10.0)
            rows += [i, j]
            cols += [j, i]
            vals += [-wgt, -wgt]
    import scipy.sparse as sp
    M = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    M.sum_duplicates()
    d = -np.asarray(M.sum(axis=1)).ravel() + rng.uniform(0.01, 1.0, size=n)
    M = (M + sp.diags(d)).tocsr()
    # shuffle the column order inside every row
    col = M.indices.copy()
    val = M.data.copy()
    for i in range(n):
        s0, s1 = M.indptr[i], M.indptr[i + 1]
        p = rng.permutation(s1 - s0)
        col[s0:s1] = col[s0:s1][p]
        val[s0:s1] = val[s0:s1][p]
    b = rng.normal(size=n)
    x_dense = np.linalg.solve(M.toarray(), b)
    for prec in (S.LS_NONE, S.LS_JACOBI):
        x, r = ls.solve(M.indptr, col, val, b, residual=1e-12, max_iterations=20000, preconditioner=prec)
        assert r.status == 0 and r.converged, (prec, r.iterations)
        assert np.abs(x - x_dense).max() <= 1e-8 * np.abs(x_dense).max(), prec
    it_none = ls.solve(M.indptr, col, val, b, residual=1e-10)[1].iterations
    it_jac = ls.solve(M.indptr, col, val, b, residual=1e-10, preconditioner=S.LS_JACOBI)[1].iterations
    say(f"  iterations plain {it_none}, jacobi {it_jac}")
    assert it_jac < it_none


def test_edge_cases(ls):
    say("linsolve edge cases")
    A = O.assemble(S.make_scene(S.SCENE_DROP, 16, seed=0))
    # zero rhs: the reference returns at once with x = 0 (pcg_solver.h:253-256)
    x, r = ls.solve(A["rowptr"], A["col"], A["val"], np.zeros(A["n"]))
    assert r.status == 0 and r.converged and r.iterations == 0 and not x.any()
    # iteration limit: not an error, converged = 0 (:292-294)
    x, r = ls.solve(A["rowptr"], A["col"], A["val"], A["b"], residual=1e-14, max_iterations=3)
    assert r.status == 0 and not r.converged and r.iterations == 3
    # NaN in the rhs
    b = A["b"].copy()
    b[5] = np.nan
    x, r = ls.solve(A["rowptr"], A["col"], A["val"], b, check=False)
    assert r.status == S.ERR_NUMERIC
    # out-of-range column
    col = A["col"].copy()
    col[0] = A["n"] + 3
    with pytest.raises(S.ShkzError):
        ls.solve(A["rowptr"], col, A["val"], A["b"])
    # empty system
    x, r = ls.solve(np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), np.zeros(0))
    assert r.converged and len(x) == 0


def test_pure_neumann_consistent_system(ls):
    """All-liquid 2D box without air: singular (constant null space); CG still converges for a
    consistent right-hand side, as the reference relies on for closed containers."""
    say("linsolve neumann")
    w = h = 24
    rng = np.random.default_rng(3)
    sc = S.Scene2D(w, h, 1 / 60, np.full((w, h), -1.0), None,
                   *(rng.normal(size=s).astype(np.float32) for s in O.face_shapes2(w, h)))
    A = O.assemble2(sc)
    b = A["b"] - A["b"].mean()
    x, r = ls.solve(A["rowptr"], A["col"], A["val"], b, residual=1e-10)
    assert r.converged
    assert np.abs(b - O.spmv(A, x)).max() <= 1e-9 * np.abs(b).max()
