"""Writes tests/golden/ref_vectors.npz and tests/golden/ref_interp_vectors.npz from the
reference's own header code.

Run in a container that holds the reference tree: it builds oracle/_ref/libshkz_ref.so
(oracle/Makefile `ref` target: the reference's include/shiokaze/utility/utility.h and
local/include/pcgsolver/sparse_matrix.h compiled where they lie) and records, as plain data,
inputs and the reference's outputs for
  utility::fraction        (include/shiokaze/utility/utility.h:162-170)
  utility::get_area        (utility.h:179-214)
  utility::box             (utility.h:113-120)
  SparseMatrix -> FixedSparseMatrix -> multiply (local/include/pcgsolver/sparse_matrix.h)
  array_interpolator3::interpolate_coef            (include/shiokaze/array/array_interpolator3.h:40-62)
  array_derivative3::derivative_interpolate_coef   (include/shiokaze/array/array_derivative3.h:40-85)
so that the tests can pin the oracle where the reference is absent.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")
OUT_INTERP = os.path.join(ROOT, "tests", "golden", "ref_interp_vectors.npz")


def special_values():
    return np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 1e-300, -1e-300, 1e-12, -1e-12, 0.25, -0.75, 3.0, -2.0],
                    np.float64)


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    R = O.ref()
    assert R is not None, "reference driver not built"
    rng = np.random.default_rng(20240601)

    # --- get_area: random quads, quads drawn from a small value set (ties, zeros, signs), sphere-like
    quads = [rng.normal(size=(6000, 4)), rng.choice(special_values(), size=(6000, 4)),
             rng.uniform(-1, 1, size=(2000, 4)) * 10.0 ** rng.integers(-8, 3, size=(2000, 1))]
    quads = np.ascontiguousarray(np.concatenate(quads), np.float64)
    area = np.array([R.ref_get_area(q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) for q in quads])

    # --- fraction
    pairs = [rng.normal(size=(6000, 2)), rng.choice(special_values(), size=(6000, 2)),
             rng.uniform(-1, 1, size=(2000, 2)) * 10.0 ** rng.integers(-300, 3, size=(2000, 1))]
    pairs = np.ascontiguousarray(np.concatenate(pairs), np.float64)
    frac = np.array([R.ref_fraction(a, b) for a, b in pairs])

    # --- box
    P = np.ascontiguousarray(rng.uniform(-0.2, 1.2, size=(2000, 3)), np.float64)
    B0 = np.ascontiguousarray(rng.uniform(0.0, 0.5, size=(2000, 3)), np.float64)
    B1 = np.ascontiguousarray(B0 + rng.uniform(0.05, 0.5, size=(2000, 3)), np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    box = np.array([R.ref_box(P[i].ctypes.data_as(dp), B0[i].ctypes.data_as(dp), B1[i].ctypes.data_as(dp))
                    for i in range(len(P))])

    # --- SparseMatrix: triplets of a small assembled pressure system, inserted in a scrambled order
    import shiokaze_amd as S
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 12, seed=5)
    sysd = O.assemble(sc)
    n = sysd["n"]
    rows = np.repeat(np.arange(n), np.diff(sysd["rowptr"]))
    perm = rng.permutation(len(rows))
    ti = np.ascontiguousarray(rows[perm], np.int64)
    tj = np.ascontiguousarray(sysd["col"][perm], np.int64)
    tv = np.ascontiguousarray(sysd["val"][perm], np.float64)
    x = np.ascontiguousarray(rng.normal(size=n), np.float64)
    y = np.zeros(n)
    rowstart = np.zeros(n + 1, np.int64)
    colindex = np.zeros(len(ti), np.int64)
    value = np.zeros(len(ti), np.float64)
    nnz = R.ref_spmv(n, len(ti), ti.ctypes.data, tj.ctypes.data, tv.ctypes.data, x.ctypes.data, y.ctypes.data,
                     rowstart.ctypes.data, colindex.ctypes.data, value.ctypes.data)
    assert nnz == len(ti)

    np.savez_compressed(OUT, quads=quads, area=area, pairs=pairs, frac=frac, box_p=P, box_p0=B0, box_p1=B1,
                        box=box, sp_n=np.int64(n), sp_ti=ti, sp_tj=tj, sp_tv=tv, sp_x=x, sp_y=y,
                        sp_rowstart=rowstart, sp_colindex=colindex, sp_value=value)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    interp_vectors(O)


def interp_vectors(O):
    """Sample positions inside, on and outside the array (the clamps), on integer and half-integer
    coordinates, for random shapes of 2..40 per axis."""
    rng = np.random.default_rng(20240602)
    m = 3000
    shapes = rng.integers(2, 41, size=(m, 3)).astype(np.int32)
    pos = np.zeros((m, 3))
    for a in range(3):
        pos[:, a] = rng.uniform(-1.5, shapes[:, a] + 0.5)
    snap = rng.integers(0, 4, size=(m, 3))  # 0 free, 1 integer, 2 half-integer, 3 the far edge
    pos = np.where(snap == 1, np.round(pos), pos)
    pos = np.where(snap == 2, np.floor(pos) + 0.5, pos)
    pos = np.where(snap == 3, shapes - 1.0, pos)
    ii = np.zeros((m, 8, 3), np.int32)
    ci = np.zeros((m, 8), np.float64)
    di = np.zeros((m, 8, 3), np.int32)
    cd = np.zeros((m, 3, 8), np.float64)
    for q in range(m):
        ii[q], ci[q] = O.ref_interp_coef(shapes[q], pos[q])
        di[q], cd[q] = O.ref_deriv_coef(shapes[q], pos[q])
    np.savez_compressed(OUT_INTERP, shapes=shapes, pos=pos, interp_idx=ii, interp_coef=ci, deriv_idx=di,
                        deriv_coef=cd)
    print("wrote", OUT_INTERP, os.path.getsize(OUT_INTERP), "bytes")


if __name__ == "__main__":
    main()
