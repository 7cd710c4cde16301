"""The Shiokaze module's staging and write-back, run by the C++ host harness (the same code,
shiokaze_amd/host/shkz_stage.hpp, that macpressuresolver3_mi355x runs), against the reference's
rules as the oracle restates them (src/projection/macpressuresolver3.cpp:245-266):

  * m_pressure holds every unknown of the system, zeros included, and nothing else (:245-248);
  * of the faces the caller's sparse velocity holds active, those the projection leaves open (and
    closed boundary faces with liquid inside) keep their projected value; the others are switched
    off; faces inactive before are not touched (velocity.parallel_actives, :252-266).

The caller's velocity is sparse: one face in five is inactive and reads its background value 0
when the module densifies it.
"""
import os
import subprocess
import sys
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "shiokaze_amd", "lib", "shkz_harness")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def read_dump(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    M, Nn = n ** 3, (n + 1) ** 3
    shapes = O.face_shapes(n, n, n)
    off = 4
    out = {"n": n}

    def take(name, dtype, shape):
        nonlocal off
        cnt = int(np.prod(shape))
        out[name] = np.frombuffer(raw[off:off + cnt * np.dtype(dtype).itemsize], dtype).reshape(shape).copy()
        off += cnt * np.dtype(dtype).itemsize

    take("fluid", np.float32, (n, n, n))
    take("solid", np.float32, (n + 1, n + 1, n + 1))
    for a in range(3):
        take(f"stage{a}", np.float32, shapes[a])
    for a in range(3):
        take(f"act_in{a}", np.uint8, shapes[a])
    for a in range(3):
        take(f"vel_after{a}", np.float32, shapes[a])
    for a in range(3):
        take(f"act_after{a}", np.uint8, shapes[a])
    take("unknown", np.uint8, (n, n, n))
    take("pressure", np.float32, (n, n, n))
    assert off == len(raw)
    assert M and Nn
    return out


@pytest.mark.parametrize("seed", [1, 5])
def test_module_writeback_follows_the_reference(tmp_path, seed):
    n = 32
    say(f"adapter write-back n={n} seed={seed}")
    assert os.path.exists(HARNESS), "build the library first (make -C shiokaze_amd/csrc)"
    path = str(tmp_path / "writeback.bin")
    subprocess.run([HARNESS, "writeback", str(n), str(seed), path], check=True, timeout=120)
    D = read_dump(path)
    sc = SimpleNamespace(w=n, h=n, d=n, dt=1.0 / 120.0, fluid=D["fluid"], solid=D["solid"],
                         u=D["stage0"].copy(), v=D["stage1"].copy(), wv=D["stage2"].copy())
    A = O.assemble(sc)
    # m_pressure: exactly the unknowns
    assert np.array_equal(D["unknown"].astype(bool), A["index_map"] >= 0)
    P = D["pressure"]
    assert (P[A["index_map"] < 0] == 0).all()
    assert A["n"] > 1000
    # the pressure solves the oracle's system (float storage of a 1e-8 solve)
    x_ref, it, rr, ok = O.pcg(A, residual=1e-12, mic_mode=O.MIC_APPLIED)
    assert ok
    x = O.gather(A, P.astype(np.float64))
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    say(f"unknowns {A['n']}, pressure rel-L2 vs the FP64 oracle {err:.2e}")
    assert err < 1e-5, err
    # the write-back equals the reference's update on the caller's sparse velocity
    act_in = [D[f"act_in{a}"] for a in range(3)]
    assert 0.15 < 1.0 - np.mean(np.concatenate([a.ravel() for a in act_in])) < 0.25
    act_ref = O.velocity_update(sc, P, A["area"], A["rho"], active=act_in, active_out=True)
    vel_ref = [sc.u, sc.v, sc.wv]
    for a in range(3):
        assert np.array_equal(D[f"act_after{a}"], act_ref[a]), a
        assert np.array_equal(D[f"vel_after{a}"].view(np.uint32), vel_ref[a].view(np.uint32)), a
        # faces inactive before stay untouched; some active faces were switched off
        assert (D[f"act_after{a}"][act_in[a] == 0] == 0).all()
        assert ((act_in[a] == 1) & (D[f"act_after{a}"] == 0)).any()
