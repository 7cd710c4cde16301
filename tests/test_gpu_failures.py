"""Host-forceable failure paths through the C ABI (the reference has none: an error prints and ends
the process, src/core/module.cpp:140-147).

* Injected device-allocation failure (shkz_inject_alloc_failure): the call that needs the failing
  allocation returns SHKZ_ERR_NOMEM with the allocation named, the caller's arrays are left as they
  were, and the same context then projects bit-identically to a fresh one.
* Request cancellation (shkz_cancel): requests cancelled while queued report SHKZ_ERR_CANCELLED and
  write none of their outputs; requests that already started run to their end; the remaining
  requests give the same pressure as direct projections; an unknown ticket is SHKZ_ERR_NOT_FOUND.
"""
import sys

import numpy as np
import pytest

import shiokaze_amd as S
from shiokaze_amd.devmem import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("after", [0, 3, 11, 29])
def test_injected_allocation_failure_leaves_context_usable(after):
    say(f"injected allocation failure after {after} allocations")
    sc = S.make_scene(S.SCENE_DROP_SOLIDS, 40, seed=1)
    prm = S.default_params(residual=1e-8)
    with S.Context(0) as fresh:
        r_ref, P_ref, P64_ref = fresh.project(ref := sc.copy(), prm, pressure64=True)
    assert r_ref.status == 0
    with S.Context(0) as ctx:
        ctx.inject_alloc_failure(after)
        work = sc.copy()
        try:
            r, P, P64 = ctx.project(work, prm, pressure64=True)
            failed = False
        except S.ShkzError as e:
            failed = True
            assert e.code == S.ERR_NOMEM and "injected" in str(e), e
        if failed:
            # nothing reached the caller's velocities
            for a, b in ((work.u, sc.u), (work.v, sc.v), (work.wv, sc.wv)):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        else:
            # fewer allocations than `after` in this call: the hook never fired; clear it
            ctx.inject_alloc_failure(-1)
        again = sc.copy()
        r2, P2, P64_2 = ctx.project(again, prm, pressure64=True)
        assert r2.status == 0 and r2.iterations == r_ref.iterations
        assert np.array_equal(P64_2, P64_ref)
        for a, b in ((again.u, ref.u), (again.v, ref.v), (again.wv, ref.wv)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert after > 0 or failed, "the first allocation of a fresh context must fail"


def test_cancel_queued_requests():
    """One executor and a queue of 12 requests; the last 6 are cancelled while the first run."""
    say("server cancellation")
    n, side = 12, 48
    prm = S.default_params(residual=1e-6)
    scenes = [S.make_scene(S.SCENE_RANDOM_DROP, side, seed=100 + i) for i in range(n)]
    st = Stream()
    dev, grids = [], []
    for sc in scenes:
        b = dict(fl=DeviceArray(sc.fluid), u=DeviceArray(sc.u), v=DeviceArray(sc.v), w=DeviceArray(sc.wv),
                 p64=DeviceArray(shape=(side, side, side), dtype=np.float64))
        b["p64"].upload(np.full((side, side, side), 7.0))  # a sentinel: a cancelled request writes nothing
        dev.append(b)
        grids.append(S.Context.device_grid(side, side, side, sc.dt, b["fl"], S.F64, None, S.F32, b["u"], b["v"],
                                           b["w"], pressure64=b["p64"]))
    try:
        with S.Server(0, executors=1, max_batch_grids=1) as srv:
            tickets = [srv.submit(g, prm) for g in grids]
            outcome = {t: srv.cancel(t) for t in reversed(tickets[n // 2:])}
            assert srv.cancel(10 ** 12) == S.ERR_NOT_FOUND
            results = [srv.wait(t) for t in tickets]
            stats = srv.stats()
        cancelled = [i for i, t in enumerate(tickets) if outcome.get(t) == S.OK]
        say(f"  cancelled {len(cancelled)} of {n // 2} asked; stats {stats}")
        assert cancelled, "no queued request could be cancelled"
        assert stats["cancelled"] == len(cancelled)
        for i, (rc, r) in enumerate(results):
            if i in cancelled:
                assert rc == S.ERR_CANCELLED and r.status == S.ERR_CANCELLED
                assert np.all(dev[i]["p64"].download() == 7.0)
                for k, a in (("u", scenes[i].u), ("v", scenes[i].v), ("w", scenes[i].wv)):
                    assert np.array_equal(dev[i][k].download().view(np.uint32), a.view(np.uint32))
            else:
                assert rc == S.OK and r.status == S.OK and r.converged, (i, rc)
        # the requests that ran equal direct projections
        with S.Context(0) as ctx:
            for i in range(n):
                if i in cancelled:
                    continue
                work = scenes[i].copy()
                r, P, P64 = ctx.project(work, prm, pressure64=True)
                assert np.array_equal(P64, dev[i]["p64"].download()), i
    finally:
        for b in dev:
            for a in b.values():
                a.free()
        st.destroy()
