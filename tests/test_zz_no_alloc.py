"""The allocation-free LinSolver entry point (shkz_linsolve_csr_device_noalloc) through the C ABI,
against the allocating entry point it varies (shkz_linsolve_csr) on the same systems: two calls in
a row, the second on a smaller system, in both preconditioner modes; results must be bit-identical
(same kernels, launch shapes and reduction order).  Also: the size limit is refused before any
launch, and an empty system returns at once."""
import ctypes
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O
from shiokaze_amd.devmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 18


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc_solve(ls, A, residual, preconditioner, stream):
    """Uploads the CSR system (32-bit indices) and solves it with the allocation-free entry."""
    L = S.lib()
    n = A["n"]
    assert 0 < n <= MAX_ROWS
    rowptr = np.ascontiguousarray(A["rowptr"], np.int32)
    col = np.ascontiguousarray(A["col"], np.int32)
    assert rowptr[0] == 0 and rowptr[-1] == len(col) and col.min() >= 0 and col.max() < n
    d = [DeviceArray(a) for a in (rowptr, col, np.ascontiguousarray(A["val"], np.float64),
                                  np.ascontiguousarray(A["b"], np.float64))]
    x = DeviceArray(shape=(n,), dtype=np.float64)
    x.zero()
    prm = S.LinsolveParams()
    L.shkz_linsolve_default_params(ctypes.byref(prm))
    prm.residual, prm.max_iterations, prm.preconditioner = residual, 30000, preconditioner
    res = S.LinsolveResult()
    rc = L.shkz_linsolve_csr_device_noalloc(ls.handle, n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                                            d[3].data_ptr(), x.data_ptr(), ctypes.byref(prm), ctypes.byref(res),
                                            stream.handle)
    assert rc == 0, S.last_error()
    stream.synchronize()
    out = x.download()
    for a in d + [x]:
        a.free()
    return out, res


@pytest.mark.parametrize("preconditioner", [S.LS_NONE, S.LS_JACOBI])
def test_noalloc_equals_allocating_entry(preconditioner):
    big = O.assemble(S.make_scene(S.SCENE_DROP_SOLIDS, 40, seed=1))
    small = O.assemble(S.make_scene(S.SCENE_DROP, 20, seed=0))
    assert big["n"] > small["n"] > 0
    stream = Stream()
    try:
        with S.LinSolver(0) as ls:
            for name, A in (("40^3 solids", big), ("20^3 drop", small)):
                say(f"noalloc linsolve {name} n={A['n']} preconditioner={preconditioner}")
                x_ref, r_ref = ls.solve(A["rowptr"], A["col"], A["val"], A["b"], residual=1e-8,
                                        preconditioner=preconditioner)
                x, r = noalloc_solve(ls, A, 1e-8, preconditioner, stream)
                assert r.status == 0 and r.converged == 1 and r_ref.converged == 1
                assert r.iterations == r_ref.iterations, (r.iterations, r_ref.iterations)
                assert r.reresid == r_ref.reresid
                assert np.array_equal(x.view(np.uint64), x_ref.view(np.uint64))
                resid = A["b"] - O.spmv(A, x)
                assert np.abs(resid).max() <= 1e-8 * np.abs(A["b"]).max() * 1.0001
    finally:
        stream.destroy()


def test_noalloc_refuses_oversize_and_handles_empty():
    L = S.lib()
    with S.LinSolver(0) as ls:
        prm = S.LinsolveParams()
        L.shkz_linsolve_default_params(ctypes.byref(prm))
        res = S.LinsolveResult()
        # over the limit: refused on the host before anything touches the (dummy) pointers
        dummy = ctypes.c_void_p(1)
        rc = L.shkz_linsolve_csr_device_noalloc(ls.handle, MAX_ROWS + 1, dummy, dummy, dummy, dummy, dummy,
                                                ctypes.byref(prm), ctypes.byref(res), None)
        assert rc == S.ERR_INVALID and "limit" in S.last_error()
        rc = L.shkz_linsolve_csr_device_noalloc(ls.handle, 0, None, None, None, None, None, ctypes.byref(prm),
                                                ctypes.byref(res), None)
        assert rc == 0 and res.converged == 1 and res.iterations == 0
