"""The multigrid transfer and neighbourhood tables (child64, parent8, nb27) are built on the device
from each level's brick map; the host build of the same tables stays behind SHKZ_HOST_LINKS=1.
Both must give bit-identical projections: one grid with solids, the 256^3 benchmark scene, a
ragged batch whose grids end on different levels, a 2D grid, and the FP64 path."""
import os
import sys

import numpy as np
import pytest

import shiokaze_amd as S

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _ctx(host_links):
    old = os.environ.get("SHKZ_HOST_LINKS")
    os.environ["SHKZ_HOST_LINKS"] = "1" if host_links else "0"
    try:
        return S.Context(0)
    finally:
        if old is None:
            del os.environ["SHKZ_HOST_LINKS"]
        else:
            os.environ["SHKZ_HOST_LINKS"] = old


@pytest.fixture(scope="module")
def ctxs():
    dev, host = _ctx(False), _ctx(True)
    yield dev, host
    dev.close()
    host.close()


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("kind,n,seed,prec", [
    (S.SCENE_DROP_SOLIDS, 64, 1, S.PREC_MIXED),
    (S.SCENE_SPHERES_BOX, 256, 2, S.PREC_MIXED),
    (S.SCENE_DROP_SOLIDS, 48, 3, S.PREC_FP64),
])
def test_device_tables_match_host_3d(ctxs, kind, n, seed, prec):
    say(f"tables 3d kind={kind} n={n} prec={prec}")
    dev, host = ctxs
    sc = S.make_scene(kind, n, seed=seed)
    prm = S.default_params(residual=1e-6, precision=prec)
    a, b = sc.copy(), sc.copy()
    ra, _, Pa = dev.project(a, prm, pressure64=True)
    rb, _, Pb = host.project(b, prm, pressure64=True)
    assert ra.status == 0 and rb.status == 0 and ra.converged
    assert ra.iterations == rb.iterations
    assert _same(Pa, Pb)
    for x, y in ((a.u, b.u), (a.v, b.v), (a.wv, b.wv)):
        assert _same(x, y)


def test_device_tables_match_host_ragged_batch(ctxs):
    say("tables ragged batch")
    dev, host = ctxs
    scenes = [S.make_scene(S.SCENE_DROP_SOLIDS, 72, 40, 24, seed=4), S.make_scene(S.SCENE_DROP, 16, seed=5),
              S.make_scene(S.SCENE_SPHERES_BOX, 96, seed=6)]
    prm = S.default_params(residual=1e-7)
    ra, _, Pa = dev.project_batch([s.copy() for s in scenes], prm, pressure64=True)
    rb, _, Pb = host.project_batch([s.copy() for s in scenes], prm, pressure64=True)
    for x, y, p, q in zip(ra, rb, Pa, Pb):
        assert x.status == 0 and y.status == 0 and x.converged
        assert x.iterations == y.iterations
        assert _same(p, q)


def test_device_tables_match_host_2d(ctxs):
    say("tables 2d")
    dev, host = ctxs
    sc = S.make_scene2d(S.SCENE_DROP_SOLIDS, 200, 136, seed=7)
    prm = S.default_params(residual=1e-8)
    a, b = sc.copy(), sc.copy()
    ra, _, Pa = dev.project2d(a, prm, pressure64=True)
    rb, _, Pb = host.project2d(b, prm, pressure64=True)
    assert ra.status == 0 and rb.status == 0 and ra.converged
    assert ra.iterations == rb.iterations
    assert _same(Pa, Pb)
    assert _same(a.u, b.u) and _same(a.v, b.v)
