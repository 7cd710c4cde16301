"""The multigrid autotuner: tunes once per shape class, reuses the cached controls, never trades
accuracy (the tuned solve converges to the requested residual) and gives the same result as an
explicit call with the chosen controls; the on-disk cache is read back by a new process."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import shiokaze_amd as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def knobs(r):
    return dict(mg_pre_sweeps=r.mg_pre_sweeps, mg_post_sweeps=r.mg_post_sweeps,
                mg_coarse_pre_sweeps=r.mg_coarse_pre_sweeps, mg_coarse_post_sweeps=r.mg_coarse_post_sweeps,
                mg_coarse_sweeps=r.mg_coarse_sweeps, mg_omega=r.mg_omega, inner_residual=r.inner_residual)


@pytest.mark.parametrize("prec", [S.PREC_MIXED, S.PREC_FP64])
def test_autotune_then_cache_then_explicit(prec):
    say(f"autotune precision {prec}")
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 72, seed=3)
    with S.Context(0) as ctx:
        r1, _, P1 = ctx.project(sc.copy(), S.default_params(residual=1e-6, precision=prec, autotune=1), pressure64=True)
        assert r1.status == 0 and r1.converged and r1.tuned == 2 and r1.reresid <= 1e-6
        r2, _, P2 = ctx.project(sc.copy(), S.default_params(residual=1e-6, precision=prec, autotune=1), pressure64=True)
        assert r2.tuned == 1 and knobs(r2) == knobs(r1)
        assert np.array_equal(P1, P2)
        r3, _, P3 = ctx.project(sc.copy(), S.default_params(residual=1e-6, precision=prec, **knobs(r1)),
                                pressure64=True)
        assert r3.tuned == 0 and knobs(r3) == knobs(r1)
        assert np.array_equal(P1, P3) and r3.iterations == r1.iterations
    say(f"  chose {knobs(r1)}")


def test_autotune_respects_fixed_controls():
    say("autotune keeps the caller's controls")
    sc = S.make_scene(S.SCENE_DROP, 48, seed=1)
    with S.Context(0) as ctx:
        r, _, _ = ctx.project(sc.copy(), S.default_params(residual=1e-6, autotune=1, mg_pre_sweeps=3,
                                                           mg_post_sweeps=3, mg_omega=0.85))
    assert r.converged and r.mg_pre_sweeps == 3 and r.mg_post_sweeps == 3 and r.mg_omega == 0.85


def test_tuning_cache_file_is_reused_by_a_new_process(tmp_path):
    say("tuning cache file")
    cache = tmp_path / "tune.txt"
    code = ("import json, shiokaze_amd as S\n"
            "sc = S.make_scene(S.SCENE_DROP, 40, seed=2)\n"
            "c = S.Context(0)\n"
            "r, _, _ = c.project(sc, S.default_params(residual=1e-6, autotune=1))\n"
            "print(json.dumps([r.tuned, r.converged, r.mg_pre_sweeps, r.mg_omega]))\n"
            "c.close()\n")
    env = dict(os.environ, SHKZ_TUNE_CACHE=str(cache), PYTHONPATH=ROOT)
    outs = []
    for _ in range(2):
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
        assert p.returncode == 0, p.stderr
        outs.append(json.loads(p.stdout.strip().splitlines()[-1]))
    assert outs[0][0] == 2 and outs[1][0] == 1
    assert outs[0][2:] == outs[1][2:] and outs[0][1] == outs[1][1] == 1
    assert len(cache.read_text().strip().splitlines()) == 1
