"""Load generator for the projection server: a served workload of time-stepped projection
requests with Poisson arrivals, driven open-loop against a shiokaze_amd.Server.

Workload (one request = one scene stepped `steps` times):
  * arrivals: exponential inter-arrival times of mean 1/rate seconds;
  * grid: sides w, h, d ~ U[side_min, side_max] (integers), redrawn until max/min <= max_aspect;
  * scene: a seeded sphere-drop (SCENE_RANDOM_DROP) whose liquid level set is rigidly translated
    by t * velocity at step t (velocity ~ U[-0.01, 0.01]^3 per step), i.e. phi_t(x) = phi_0(x - t v);
  * fresh N(0,1) face velocities every step (counter-based device generator);
  * warm start from the scene's previous pressure, which stays resident in the server.
Inputs are generated on the device (shkz_scene_fill_device) on the client's stream and are complete
before each submission.  A request's next step is submitted when its previous step completes.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import (F32, PENDING, SCENE_RANDOM_DROP, Context, ShkzError, default_params, last_error, make_scene,
               scene_fill_device)
from .devmem import Stream, check, hip

# Batched workload: 64 cubes with sides drawn uniformly from these (numpy default_rng(3)).
C4_SIDES = (32, 48, 64, 96, 128, 160)
C4_GRIDS = 64


def c4_sides(n_grids: int = C4_GRIDS, seed: int = 3) -> list[int]:
    return [int(a) for a in np.random.default_rng(seed).choice(C4_SIDES, size=n_grids)]


def c4_scenes(n_grids: int = C4_GRIDS, seed: int = 3, workers: int = 16) -> list:
    """The batched workload's host scenes: cube i has side c4_sides()[i] and is a seeded
    sphere-drop (scene seed 1000*seed + i) with its own N(0,1) velocities."""
    from concurrent.futures import ThreadPoolExecutor
    sides = c4_sides(n_grids, seed)
    # the generator runs in the library (ctypes releases the GIL): synthesise in parallel
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(lambda a: make_scene(SCENE_RANDOM_DROP, a[1], seed=1000 * seed + a[0]),
                           enumerate(sides)))


@dataclass
class RequestSpec:
    idx: int
    t_arrival: float  # seconds from the start of the run
    w: int
    h: int
    d: int
    seed: int
    velocity: tuple   # level-set translation per step


def make_requests(n: int, rate: float = 200.0, seed: int = 4, side_min: int = 48, side_max: int = 192,
                  max_aspect: float = 2.0) -> list[RequestSpec]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        while True:
            s = rng.integers(side_min, side_max + 1, size=3)
            if s.max() <= max_aspect * s.min():
                break
        vel = tuple(float(x) for x in rng.uniform(-0.01, 0.01, size=3))
        out.append(RequestSpec(i, t, int(s[0]), int(s[1]), int(s[2]), seed * 100000 + i, vel))
    return out


class _Pool:
    """Device buffers reused across requests (best fit by capacity, never freed during a run:
    hipFree would synchronise the device under the server's work)."""

    def __init__(self):
        self.free: list[tuple[int, int]] = []  # (capacity, ptr)
        self.all: list[int] = []

    def take(self, nbytes: int) -> tuple[int, int]:
        best = None
        for i, (cap, p) in enumerate(self.free):
            if cap >= nbytes and (best is None or cap < self.free[best][0]):
                best = i
        if best is not None:
            return self.free.pop(best)
        cap = int(nbytes * 1.25) + (1 << 20)
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), cap), "hipMalloc")
        self.all.append(p.value)
        return cap, p.value

    def give(self, block: tuple[int, int]):
        self.free.append(block)

    def close(self):
        for p in self.all:
            hip().hipFree(p)
        self.all.clear()
        self.free.clear()


def _align(n: int) -> int:
    return (n + 255) & ~255


@dataclass
class _Active:
    spec: RequestSpec
    block: tuple
    ptrs: dict
    step: int = 0
    ticket: int = 0
    t_start: float = 0.0
    solve_lat: list = field(default_factory=list)
    iters: list = field(default_factory=list)
    pressure64: object = None


def step_inputs(spec: RequestSpec, step: int, ptrs: dict, stream) -> None:
    shift = tuple(step * v for v in spec.velocity)
    scene_fill_device(SCENE_RANDOM_DROP, spec.w, spec.h, spec.d, ptrs["fluid"], ptrs["u"], ptrs["v"], ptrs["w"],
                      seed=spec.seed, fluid_dtype=F32, shift=shift, velocity_seed=spec.seed * 64 + step + 1,
                      stream=stream)


def layout(spec: RequestSpec) -> tuple[int, dict]:
    w, h, d = spec.w, spec.h, spec.d
    sizes = {"fluid": 4 * w * h * d, "u": 4 * (w + 1) * h * d, "v": 4 * w * (h + 1) * d, "w": 4 * w * h * (d + 1)}
    off, o = {}, 0
    for k, n in sizes.items():
        off[k] = o
        o += _align(n)
    return o, off


def run_served(server, specs: list[RequestSpec], steps: int = 20, params=None, time_scale: float = 1.0,
               keep_pressure: set | None = None, log=None) -> dict:
    """Drives `server` open-loop with `specs` (arrival times scaled by time_scale).  Returns the
    per-request and per-solve latencies (ms), iteration counts, statuses and the wall time."""
    params = params if params is not None else default_params(residual=1e-6, warm_start=1)
    st = Stream()
    pool = _Pool()
    pending = sorted(specs, key=lambda s: s.t_arrival)
    active: dict[int, _Active] = {}
    req_lat, solve_lat, iters, statuses = [], [], [], []
    pressures = {}
    solves = 0
    t0 = time.perf_counter()
    next_log = 5.0
    pi = 0
    try:
        while pi < len(pending) or active:
            now = time.perf_counter() - t0
            progressed = False
            while pi < len(pending) and pending[pi].t_arrival * time_scale <= now:
                spec = pending[pi]
                pi += 1
                nbytes, off = layout(spec)
                block = pool.take(nbytes)
                ptrs = {k: block[1] + o for k, o in off.items()}
                a = _Active(spec, block, ptrs, t_start=now)
                if keep_pressure and spec.idx in keep_pressure:
                    from .devmem import DeviceArray
                    a.pressure64 = DeviceArray(shape=(steps, spec.w, spec.h, spec.d), dtype=np.float64)
                _submit_step(server, a, params, st)
                active[spec.idx] = a
                progressed = True
            for idx in list(active):
                a = active[idx]
                rc, r = server.poll(a.ticket)
                if rc == PENDING:
                    continue
                progressed = True
                statuses.append(rc)
                solves += 1
                solve_lat.append(r.t_latency)
                iters.append(r.iterations)
                a.step += 1
                if rc == 0 and a.step < steps:
                    _submit_step(server, a, params, st)
                    continue
                t_done = time.perf_counter() - t0
                req_lat.append((t_done - a.spec.t_arrival * time_scale) * 1e3)
                try:
                    server.release_scene(a.spec.idx + 1)
                except ShkzError:
                    pass  # a request that failed before its scene was created
                if a.pressure64 is not None:
                    pressures[idx] = a.pressure64.download()
                    a.pressure64.free()
                pool.give(a.block)
                del active[idx]
            if log and now > next_log:
                log(f"[serve] t={now:.1f}s arrived {pi}/{len(pending)} active {len(active)} solves {solves}")
                next_log += 5.0
            if not progressed:
                time.sleep(0.0002)
        elapsed = time.perf_counter() - t0
    finally:
        st.synchronize()
        pool.close()
        st.destroy()
    return {"request_latency_ms": req_lat, "solve_latency_ms": solve_lat, "iterations": iters,
            "statuses": statuses, "solves": solves, "elapsed_s": elapsed, "pressures": pressures}


def _submit_step(server, a: _Active, params, st):
    step_inputs(a.spec, a.step, a.ptrs, st.handle)
    st.synchronize()  # inputs complete before submission
    p64 = None
    if a.pressure64 is not None:
        p64 = a.pressure64.data_ptr() + a.step * 8 * a.spec.w * a.spec.h * a.spec.d
    g = Context.device_grid(a.spec.w, a.spec.h, a.spec.d, 1.0 / 120.0, a.ptrs["fluid"], F32, None, F32,
                            a.ptrs["u"], a.ptrs["v"], a.ptrs["w"], pressure64=p64, scene_id=a.spec.idx + 1)
    try:
        a.ticket = server.submit(g, params)
    except ShkzError:
        raise RuntimeError(f"submit failed: {last_error()}")
