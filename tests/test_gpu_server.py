"""The projection server (continuous batching): each request's result is the same alone and in
the mix, and equals the direct shkz_project_device result; admission by the device budget; device
scene synthesis matches the host generator's level sets."""
import dataclasses
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from shiokaze_amd import loadgen
from shiokaze_amd.devmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind,dims,shift", [(S.SCENE_DROP_SOLIDS, (40, 40, 40), (0, 0, 0)),
                                             (S.SCENE_SPHERES_BOX, (33, 47, 29), (0, 0, 0)),
                                             (S.SCENE_RANDOM_DROP, (48, 40, 56), (0.013, -0.02, 0.004))])
def test_device_scene_fill_matches_host_level_sets(kind, dims, shift):
    say(f"device scene fill kind {kind} {dims}")
    w, h, d = dims
    host = S.make_scene(kind, w, h, d, seed=11, shift=shift)
    fl = DeviceArray(shape=(w, h, d), dtype=np.float64)
    so = DeviceArray(shape=(w + 1, h + 1, d + 1), dtype=np.float32)
    u = DeviceArray(shape=(w + 1, h, d), dtype=np.float32)
    v = DeviceArray(shape=(w, h + 1, d), dtype=np.float32)
    wv = DeviceArray(shape=(w, h, d + 1), dtype=np.float32)
    st = Stream()
    has = S.scene_fill_device(kind, w, h, d, fl, u, v, wv, seed=11, shift=shift, solid=so, stream=st.handle)
    st.synchronize()
    assert has == (host.solid is not None)
    assert np.array_equal(fl.download(), host.fluid)
    if host.solid is not None:
        assert np.array_equal(so.download(), host.solid)
    vel = np.concatenate([a.download().ravel() for a in (u, wv)])
    assert abs(vel.mean()) < 0.02 and abs(vel.std() - 1.0) < 0.02
    vv = v.download()
    assert abs(vv.mean() - (-9.8 / 120.0)) < 0.03
    for a in (fl, so, u, v, wv):
        a.free()
    st.destroy()


def specs_mix(steps):
    probe = loadgen.RequestSpec(0, 0.0, 40, 56, 48, 424242, (0.004, -0.006, 0.002))
    others = loadgen.make_requests(12, rate=1e6, seed=9, side_min=32, side_max=96)
    others = [dataclasses.replace(s, idx=s.idx + 1) for s in others]
    return probe, others


def test_server_alone_vs_in_the_mix_vs_direct():
    steps = 3
    probe, others = specs_mix(steps)
    prm = S.default_params(residual=1e-6, warm_start=1)
    say("server: probe alone")
    with S.Server(0, executors=2) as srv:
        alone = loadgen.run_served(srv, [probe], steps=steps, params=prm, keep_pressure={0})
    say("server: probe in a mix of 13 requests")
    with S.Server(0, executors=2, max_batch_grids=8) as srv:
        mix = loadgen.run_served(srv, [probe] + others, steps=steps, params=prm, keep_pressure={0})
        st = srv.stats()
    say(f"  stats {st}")
    assert all(s == 0 for s in alone["statuses"]) and all(s == 0 for s in mix["statuses"])
    assert mix["solves"] == 13 * steps
    assert st["batches"] < st["batched_requests"], "no request shared a batch"
    assert np.array_equal(alone["pressures"][0], mix["pressures"][0])
    # the same steps through shkz_project_device on one context with the same scene id
    say("direct: probe stepped with shkz_project_device")
    nbytes, off = loadgen.layout(probe)
    buf = DeviceArray(shape=(nbytes,), dtype=np.uint8)
    ptrs = {k: buf.data_ptr() + o for k, o in off.items()}
    p64 = DeviceArray(shape=(probe.w, probe.h, probe.d), dtype=np.float64)
    stream = Stream()
    ctx = S.Context(0)
    try:
        for t in range(steps):
            loadgen.step_inputs(probe, t, ptrs, stream.handle)
            r = ctx.project_device(probe.w, probe.h, probe.d, 1.0 / 120.0, ptrs["fluid"], S.F32, None, S.F32,
                                   ptrs["u"], ptrs["v"], ptrs["w"], params=prm, pressure64=p64, scene_id=1,
                                   stream=stream.handle)
            stream.synchronize()
            assert r.status == 0
            assert np.array_equal(p64.download(), alone["pressures"][0][t]), t
    finally:
        ctx.close()
        buf.free()
        p64.free()
        stream.destroy()


def test_server_admission_by_budget():
    say("server budget admission")
    small = loadgen.RequestSpec(0, 0.0, 32, 32, 32, 5, (0.0, 0.0, 0.0))
    big = loadgen.RequestSpec(1, 0.0, 160, 160, 160, 6, (0.0, 0.0, 0.0))
    est_big = 160 ** 3 * 100
    with S.Server(0, executors=2, hbm_budget_bytes=est_big // 2) as srv:
        res = loadgen.run_served(srv, [small, big], steps=1, params=S.default_params(residual=1e-6))
        st = srv.stats()
    assert sorted(res["statuses"]) == [S.ERR_BUDGET, 0]
    assert st["rejected"] == 1
