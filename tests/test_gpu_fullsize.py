"""The projection at the benchmark configurations' full sizes (128^3 with solids, 256^3), through
properties that do not depend on size:

  * the FP64 residual of the returned pressure, measured on the oracle's own assembly of the
    system (CSR, FP64), meets the requested tolerance (infinity norm relative to b);
  * the returned float pressure is the FP64 pressure rounded;
  * the updated velocities are bit-identical to the oracle's velocity update applied to the GPU's
    pressure (every face of the box);
  * a second projection of the same input gives bit-identical results (determinism at scale).

The oracle assembles and applies the operator here; it does not solve (a CPU solve at 256^3 takes
minutes).
"""
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FULL = {
    # name: (scene kind, n, seed, residual) as the benchmark configurations set them
    "c2_128": (S.SCENE_DROP_SOLIDS, 128, 1, 1e-8),
    "c3_256": (S.SCENE_SPHERES_BOX, 256, 2, 1e-6),
}


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", list(FULL))
def test_full_size_projection(ctx, name):
    kind, n, seed, residual = FULL[name]
    say(f"full size {name}")
    sc = S.make_scene(kind, n, seed=seed)
    prm = S.default_params(residual=residual, precision=S.PREC_MIXED)
    work = sc.copy()
    r, P, P64 = ctx.project(work, prm, pressure64=True)
    assert r.status == 0 and r.converged, r.as_dict()
    assert np.array_equal(P, P64.astype(np.float32))

    A = O.assemble(sc)
    assert r.n_unknowns == A["n"]
    x = O.gather(A, P64)
    resid = A["b"] - O.spmv(A, x)
    rel = np.abs(resid).max() / np.abs(A["b"]).max()
    say(f"{name}: n={A['n']} iterations={r.iterations} residual={rel:.3e}")
    assert rel <= residual * 1.0001, rel

    ref = sc.copy()
    O.velocity_update(ref, P, A["area"], A["rho"])
    for a, b in ((work.u, ref.u), (work.v, ref.v), (work.wv, ref.wv)):
        assert np.array_equal(a, b)

    again = sc.copy()
    r2, P2, P64b = ctx.project(again, prm, pressure64=True)
    assert r2.iterations == r.iterations
    assert np.array_equal(P64b, P64)
    for a, b in ((again.u, work.u), (again.v, work.v), (again.wv, work.wv)):
        assert np.array_equal(a, b)
