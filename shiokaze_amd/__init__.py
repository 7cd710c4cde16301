"""shiokaze_amd -- MI355X pressure projection for Shiokaze, Python binding of the C ABI.

The product is ``shiokaze_amd/lib/libshkz_mi355x.so`` (C ABI declared in
``include/shkz_mi355x.h``); this module only binds it with ctypes for tests and the
benchmark.  It performs no computation of its own and has no fallback: if the library is
missing, importing works but every call raises ``RuntimeError``.

``Context.project`` mirrors ``macproject3_interface::project``
(include/shiokaze/projection/macproject3_interface.h:67-72) on dense numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libshkz_mi355x.so")

ABI_VERSION = 8  # the layout of the structs below
F32, F64 = 0, 1
SOLVER_MGPCG, SOLVER_CG = 0, 1
PREC_MIXED, PREC_FP64 = 0, 1
NORM_INF, NORM_L2 = 0, 1
SCENE_DROP, SCENE_DROP_SOLIDS, SCENE_SPHERES_BOX, SCENE_RANDOM_DROP, SCENE_SMOKE = 0, 1, 2, 3, 4

OK, ERR_INVALID, ERR_DEVICE, ERR_NOMEM, ERR_BUDGET, ERR_NOT_FOUND, ERR_NUMERIC = 0, -1, -2, -3, -4, -5, -6
ERR_CANCELLED = -7
STATUS = {0: "OK", -1: "INVALID", -2: "DEVICE", -3: "NOMEM", -4: "BUDGET", -5: "NOT_FOUND", -6: "NUMERIC",
          -7: "CANCELLED"}

_vp = ctypes.c_void_p


class Grid(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("d", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("fluid", _vp), ("fluid_dtype", ctypes.c_int32), ("solid_dtype", ctypes.c_int32),
        ("solid", _vp),
        ("u", _vp), ("v", _vp), ("w_vel", _vp),
        ("active_u", _vp), ("active_v", _vp), ("active_w", _vp),
        ("pressure", _vp), ("pressure64", _vp),
        ("scene_id", ctypes.c_uint64),
        ("active_out_u", _vp), ("active_out_v", _vp), ("active_out_w", _vp),
        ("unknown_out", _vp),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("residual", ctypes.c_double), ("norm", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
        ("solver", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("eps_fluid", ctypes.c_double), ("eps_solid", ctypes.c_double),
        ("second_order_fluid", ctypes.c_int32), ("second_order_solid", ctypes.c_int32),
        ("surface_tension", ctypes.c_double), ("gain", ctypes.c_double),
        ("current_volume", ctypes.c_double), ("target_volume", ctypes.c_double),
        ("warm_start", ctypes.c_int32),
        ("mg_pre_sweeps", ctypes.c_int32), ("mg_post_sweeps", ctypes.c_int32), ("mg_coarse_sweeps", ctypes.c_int32),
        ("mg_omega", ctypes.c_double), ("inner_residual", ctypes.c_double),
        ("mg_coarse_pre_sweeps", ctypes.c_int32), ("mg_coarse_post_sweeps", ctypes.c_int32),
        ("mg_bottom_bricks", ctypes.c_int32), ("autotune", ctypes.c_int32),
        ("mg_smoother", ctypes.c_int32), ("mg_local_sweeps", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("converged", ctypes.c_int32), ("iterations", ctypes.c_int32),
        ("outer_iterations", ctypes.c_int32),
        ("reresid", ctypes.c_double), ("reresid_inf", ctypes.c_double), ("reresid_l2", ctypes.c_double),
        ("n_unknowns", ctypes.c_int64), ("n_bricks", ctypes.c_int64), ("n_levels", ctypes.c_int32),
        ("batch_size", ctypes.c_int32), ("rhs_correct", ctypes.c_double),
        ("t_solid_fluid_fractions", ctypes.c_double), ("t_build_highres_linsystem", ctypes.c_double),
        ("t_linsolve", ctypes.c_double), ("t_update_velocity", ctypes.c_double), ("t_projection", ctypes.c_double),
        ("t_queue", ctypes.c_double), ("t_latency", ctypes.c_double),
        ("mg_pre_sweeps", ctypes.c_int32), ("mg_post_sweeps", ctypes.c_int32),
        ("mg_coarse_pre_sweeps", ctypes.c_int32), ("mg_coarse_post_sweeps", ctypes.c_int32),
        ("mg_coarse_sweeps", ctypes.c_int32), ("tuned", ctypes.c_int32),
        ("mg_omega", ctypes.c_double), ("inner_residual", ctypes.c_double),
        ("mg_smoother", ctypes.c_int32), ("mg_local_sweeps", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class VelocityField(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("d", ctypes.c_int32), ("solid_dtype", ctypes.c_int32),
        ("dx", ctypes.c_double), ("solid", _vp),
        ("u", _vp), ("v", _vp), ("w_vel", _vp),
        ("active_u", _vp), ("active_v", _vp), ("active_w", _vp),
    ]


class AdvectParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("maccormack", ctypes.c_int32), ("trim_narrowband", ctypes.c_int32),
                ("fluid", _vp), ("fluid_dtype", ctypes.c_int32), ("background", ctypes.c_float),
                ("weno", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class Grid2D(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("fluid", _vp), ("fluid_dtype", ctypes.c_int32), ("solid_dtype", ctypes.c_int32), ("solid", _vp),
        ("u", _vp), ("v", _vp), ("active_u", _vp), ("active_v", _vp), ("pressure", _vp), ("pressure64", _vp),
        ("scene_id", ctypes.c_uint64), ("unknown_out", _vp),
    ]


class LinsolveParams(ctypes.Structure):
    _fields_ = [("residual", ctypes.c_double), ("max_iterations", ctypes.c_int32), ("preconditioner", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 6)]


class LinsolveResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("converged", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("reresid", ctypes.c_double), ("t_upload", ctypes.c_double),
                ("t_solve", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("executors", ctypes.c_int32), ("max_batch_grids", ctypes.c_int32),
        ("reserved0", ctypes.c_int32), ("max_batch_cells", ctypes.c_int64), ("hbm_budget_bytes", ctypes.c_uint64),
        ("reserved", ctypes.c_int32 * 8),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
        ("batches", ctypes.c_uint64), ("batched_requests", ctypes.c_uint64), ("max_batch", ctypes.c_int32),
        ("max_in_flight_batches", ctypes.c_int32), ("max_in_flight_bytes", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64), ("cancelled", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("d", ctypes.c_int32),
        ("seed", ctypes.c_uint64), ("dt", ctypes.c_double), ("gravity", ctypes.c_double),
        ("shift", ctypes.c_double * 3), ("fluid_dtype", ctypes.c_int32), ("velocity_mode", ctypes.c_int32),
        ("velocity_seed", ctypes.c_uint64),
    ]


# Every symbol declared in include/shkz_mi355x.h.
EXPORTS = [
    "shkz_abi_version", "shkz_last_error", "shkz_default_params", "shkz_create", "shkz_destroy",
    "shkz_project", "shkz_project_device", "shkz_release_scene", "shkz_device_bytes",
    "shkz_set_kernel_timing", "shkz_get_kernel_stats", "shkz_reset_kernel_stats", "shkz_scene_fill",
    "shkz_export_system", "shkz_project_batch", "shkz_project_batch_device",
    "shkz_server_default_config", "shkz_server_create", "shkz_server_destroy", "shkz_submit", "shkz_poll",
    "shkz_wait", "shkz_server_release_scene", "shkz_server_get_stats", "shkz_scene_fill_device",
    "shkz_project2d", "shkz_project2d_device", "shkz_project2d_batch_device", "shkz_export_system2d",
    "shkz_scene_fill2d", "shkz_linsolve_default_params", "shkz_linsolver_create", "shkz_linsolver_destroy",
    "shkz_linsolve_csr", "shkz_extrapolate_velocity", "shkz_extrapolate_velocity_device",
    "shkz_advect_default_params", "shkz_advect_velocity", "shkz_advect_velocity_device", "shkz_advect_scalar",
    "shkz_advect_scalar_device", "shkz_linsolve_csr_device_noalloc", "shkz_cancel", "shkz_inject_alloc_failure",
]

_lib = None


def lib():
    """Loads libshkz_mi355x.so (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C shiokaze_amd/csrc); there is no fallback implementation")
    L = ctypes.CDLL(LIB_PATH)
    L.shkz_abi_version.restype = ctypes.c_int
    if L.shkz_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI {L.shkz_abi_version()}, these bindings expect {ABI_VERSION}: rebuild it")
    L.shkz_last_error.restype = ctypes.c_char_p
    L.shkz_default_params.argtypes = [ctypes.POINTER(Params)]
    L.shkz_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]
    L.shkz_destroy.argtypes = [_vp]
    L.shkz_project.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.POINTER(Result)]
    L.shkz_project_device.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.POINTER(Result), _vp]
    L.shkz_release_scene.argtypes = [_vp, ctypes.c_uint64]
    L.shkz_inject_alloc_failure.argtypes = [_vp, ctypes.c_int32]
    L.shkz_cancel.argtypes = [_vp, ctypes.c_uint64]
    L.shkz_device_bytes.argtypes = [_vp]
    L.shkz_device_bytes.restype = ctypes.c_uint64
    L.shkz_set_kernel_timing.argtypes = [_vp, ctypes.c_int]
    L.shkz_get_kernel_stats.argtypes = [_vp, ctypes.POINTER(KernelStat), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.shkz_reset_kernel_stats.argtypes = [_vp]
    L.shkz_scene_fill.argtypes = [ctypes.POINTER(SceneDesc), _vp, _vp, ctypes.POINTER(ctypes.c_int32), _vp, _vp, _vp]
    L.shkz_export_system.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.POINTER(Params), _vp, _vp, _vp, _vp, _vp]
    L.shkz_project_batch.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.c_int, ctypes.POINTER(Params),
                                     ctypes.POINTER(Result)]
    L.shkz_project_batch_device.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.c_int, ctypes.POINTER(Params),
                                            ctypes.POINTER(Result), _vp]
    L.shkz_server_default_config.argtypes = [ctypes.POINTER(ServerConfig)]
    L.shkz_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.shkz_server_destroy.argtypes = [_vp]
    L.shkz_submit.argtypes = [_vp, ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_uint64)]
    L.shkz_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(Result)]
    L.shkz_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(Result)]
    L.shkz_server_release_scene.argtypes = [_vp, ctypes.c_uint64]
    L.shkz_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    L.shkz_scene_fill_device.argtypes = [ctypes.POINTER(SceneDesc), _vp, _vp, ctypes.POINTER(ctypes.c_int32), _vp,
                                         _vp, _vp, _vp]
    L.shkz_project2d.argtypes = [_vp, ctypes.POINTER(Grid2D), ctypes.POINTER(Params), ctypes.POINTER(Result)]
    L.shkz_project2d_device.argtypes = [_vp, ctypes.POINTER(Grid2D), ctypes.POINTER(Params), ctypes.POINTER(Result),
                                        _vp]
    L.shkz_project2d_batch_device.argtypes = [_vp, ctypes.POINTER(Grid2D), ctypes.c_int, ctypes.POINTER(Params),
                                              ctypes.POINTER(Result), _vp]
    L.shkz_export_system2d.argtypes = [_vp, ctypes.POINTER(Grid2D), ctypes.POINTER(Params), _vp, _vp, _vp, _vp]
    L.shkz_scene_fill2d.argtypes = [ctypes.POINTER(SceneDesc), _vp, _vp, ctypes.POINTER(ctypes.c_int32), _vp, _vp]
    L.shkz_linsolve_default_params.argtypes = [ctypes.POINTER(LinsolveParams)]
    L.shkz_linsolver_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.shkz_linsolver_destroy.argtypes = [_vp]
    L.shkz_linsolve_csr.argtypes = [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(LinsolveParams),
                                    ctypes.POINTER(LinsolveResult)]
    L.shkz_linsolve_csr_device_noalloc.argtypes = [_vp, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp,
                                                   ctypes.POINTER(LinsolveParams), ctypes.POINTER(LinsolveResult), _vp]
    L.shkz_extrapolate_velocity.argtypes = [_vp, ctypes.POINTER(VelocityField), ctypes.c_int32, ctypes.c_int32]
    L.shkz_extrapolate_velocity_device.argtypes = [_vp, ctypes.POINTER(VelocityField), ctypes.c_int32, ctypes.c_int32,
                                                   _vp]
    L.shkz_advect_default_params.argtypes = [ctypes.POINTER(AdvectParams)]
    L.shkz_advect_velocity.argtypes = [_vp, ctypes.POINTER(VelocityField), ctypes.POINTER(AdvectParams)]
    L.shkz_advect_velocity_device.argtypes = [_vp, ctypes.POINTER(VelocityField), ctypes.POINTER(AdvectParams), _vp]
    L.shkz_advect_scalar.argtypes = [_vp, ctypes.POINTER(VelocityField), _vp, _vp, ctypes.POINTER(AdvectParams)]
    L.shkz_advect_scalar_device.argtypes = [_vp, ctypes.POINTER(VelocityField), _vp, _vp, ctypes.POINTER(AdvectParams),
                                            _vp]
    _lib = L
    return L


def _dp(x):
    """Device pointer of an int or an object with data_ptr() (None stays None)."""
    if x is None:
        return None
    return x.data_ptr() if hasattr(x, "data_ptr") else int(x)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return int(a)  # device pointer as an int


def last_error() -> str:
    return lib().shkz_last_error().decode()


class ShkzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def default_params(**kw) -> Params:
    p = Params()
    lib().shkz_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


@dataclass
class Scene:
    """Dense host arrays of one projection request (layouts of include/shkz_mi355x.h)."""
    w: int
    h: int
    d: int
    dt: float
    fluid: np.ndarray
    solid: np.ndarray | None
    u: np.ndarray
    v: np.ndarray
    wv: np.ndarray

    @property
    def dx(self) -> float:
        return min(1.0 / self.w, 1.0 / self.h, 1.0 / self.d)

    def copy(self) -> "Scene":
        return Scene(self.w, self.h, self.d, self.dt, self.fluid.copy(),
                     None if self.solid is None else self.solid.copy(), self.u.copy(), self.v.copy(), self.wv.copy())


def make_scene(kind: int, w: int, h: int | None = None, d: int | None = None, seed: int = 0, dt: float = 1.0 / 120.0,
               gravity: float = -9.8, fluid_dtype: int = F64, shift=(0.0, 0.0, 0.0), velocity_mode: int = 0,
               velocity_seed: int = 0) -> Scene:
    """Synthesises a seeded workload with shkz_scene_fill."""
    h = w if h is None else h
    d = w if d is None else d
    desc = SceneDesc(kind=kind, w=w, h=h, d=d, seed=seed, dt=dt, gravity=gravity, fluid_dtype=fluid_dtype,
                     velocity_mode=velocity_mode, velocity_seed=velocity_seed)
    for a in range(3):
        desc.shift[a] = shift[a]
    fluid = np.empty((w, h, d), np.float64 if fluid_dtype == F64 else np.float32)
    solid = np.empty((w + 1, h + 1, d + 1), np.float32)
    u = np.empty((w + 1, h, d), np.float32)
    v = np.empty((w, h + 1, d), np.float32)
    wv = np.empty((w, h, d + 1), np.float32)
    has = ctypes.c_int32(0)
    rc = lib().shkz_scene_fill(ctypes.byref(desc), _ptr(fluid), _ptr(solid), ctypes.byref(has), _ptr(u), _ptr(v),
                               _ptr(wv))
    if rc != 0:
        raise ShkzError(rc, "shkz_scene_fill failed")
    return Scene(w, h, d, dt, fluid, solid if has.value else None, u, v, wv)


@dataclass
class Scene2D:
    """Dense host arrays of one 2D projection request (y fastest): fluid (w, h), solid (w+1, h+1),
    u (w+1, h), v (w, h+1)."""
    w: int
    h: int
    dt: float
    fluid: np.ndarray
    solid: np.ndarray | None
    u: np.ndarray
    v: np.ndarray

    @property
    def dx(self) -> float:
        return min(1.0 / self.w, 1.0 / self.h)

    def copy(self) -> "Scene2D":
        return Scene2D(self.w, self.h, self.dt, self.fluid.copy(), None if self.solid is None else self.solid.copy(),
                       self.u.copy(), self.v.copy())


def make_scene2d(kind: int, w: int, h: int | None = None, seed: int = 0, dt: float = 1.0 / 120.0,
                 gravity: float = -9.8, fluid_dtype: int = F64, shift=(0.0, 0.0), velocity_mode: int = 0,
                 velocity_seed: int = 0) -> Scene2D:
    """Synthesises a seeded 2D workload with shkz_scene_fill2d."""
    h = w if h is None else h
    desc = SceneDesc(kind=kind, w=w, h=h, d=1, seed=seed, dt=dt, gravity=gravity, fluid_dtype=fluid_dtype,
                     velocity_mode=velocity_mode, velocity_seed=velocity_seed)
    desc.shift[0], desc.shift[1] = shift[0], shift[1]
    fluid = np.empty((w, h), np.float64 if fluid_dtype == F64 else np.float32)
    solid = np.empty((w + 1, h + 1), np.float32)
    u = np.empty((w + 1, h), np.float32)
    v = np.empty((w, h + 1), np.float32)
    has = ctypes.c_int32(0)
    rc = lib().shkz_scene_fill2d(ctypes.byref(desc), _ptr(fluid), _ptr(solid), ctypes.byref(has), _ptr(u), _ptr(v))
    if rc != 0:
        raise ShkzError(rc, "shkz_scene_fill2d failed")
    return Scene2D(w, h, dt, fluid, solid if has.value else None, u, v)


def scene_fill_device(kind: int, w: int, h: int, d: int, fluid, u, v, wv, seed: int = 0, dt: float = 1.0 / 120.0,
                      gravity: float = -9.8, fluid_dtype: int = F64, shift=(0.0, 0.0, 0.0), velocity_mode: int = 0,
                      velocity_seed: int = 0, solid=None, stream=None) -> bool:
    """Synthesises a scene into device buffers (objects with data_ptr() or ints) on `stream` with
    shkz_scene_fill_device; returns whether the scene has a solid."""
    def dp(x):
        if x is None:
            return None
        return x.data_ptr() if hasattr(x, "data_ptr") else int(x)
    desc = SceneDesc(kind=kind, w=w, h=h, d=d, seed=seed, dt=dt, gravity=gravity, fluid_dtype=fluid_dtype,
                     velocity_mode=velocity_mode, velocity_seed=velocity_seed)
    for a in range(3):
        desc.shift[a] = shift[a]
    has = ctypes.c_int32(0)
    rc = lib().shkz_scene_fill_device(ctypes.byref(desc), dp(fluid), dp(solid), ctypes.byref(has), dp(u), dp(v),
                                      dp(wv), None if stream is None else int(stream))
    if rc != 0:
        raise ShkzError(rc, "shkz_scene_fill_device failed")
    return bool(has.value)


PENDING = 1
LS_NONE, LS_JACOBI = 0, 1
SMOOTH_DEFAULT, SMOOTH_JACOBI, SMOOTH_BRICK = 0, 1, 2


class LinSolver:
    """The LinSolver drop-in (shkz_linsolver): CG / Jacobi-PCG on a CSR system in FP64."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = _vp()
        rc = self._L.shkz_linsolver_create(device, ctypes.byref(h))
        if rc != 0:
            raise ShkzError(rc, last_error())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._L.shkz_linsolver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, rowptr, col, val, b, residual=1e-4, max_iterations=30000, preconditioner=LS_NONE,
              check=True):
        """Solves A x = b (CSR, int64 indices, float64 values).  Returns (x, result)."""
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col = np.ascontiguousarray(col, np.int64)
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = len(rowptr) - 1
        x = np.zeros(max(n, 1), np.float64)
        prm = LinsolveParams()
        self._L.shkz_linsolve_default_params(ctypes.byref(prm))
        prm.residual, prm.max_iterations, prm.preconditioner = residual, max_iterations, preconditioner
        res = LinsolveResult()
        rc = self._L.shkz_linsolve_csr(self.handle, n, rowptr.ctypes.data, col.ctypes.data if len(col) else None,
                                       val.ctypes.data if len(val) else None, b.ctypes.data, x.ctypes.data,
                                       ctypes.byref(prm), ctypes.byref(res))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return x[:n], res


class Server:
    """A projection server (shkz_server): continuous batching of submitted device-resident requests."""

    def __init__(self, device: int = 0, executors: int = 2, max_batch_grids: int = 64, max_batch_cells: int = 0,
                 hbm_budget_bytes: int = 0):
        self._L = lib()
        cfg = ServerConfig()
        self._L.shkz_server_default_config(ctypes.byref(cfg))
        cfg.device = device
        cfg.executors = executors
        cfg.max_batch_grids = max_batch_grids
        if max_batch_cells:
            cfg.max_batch_cells = max_batch_cells
        cfg.hbm_budget_bytes = hbm_budget_bytes
        h = _vp()
        rc = self._L.shkz_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise ShkzError(rc, last_error())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._L.shkz_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, grid: "Grid", params: "Params | None" = None) -> int:
        t = ctypes.c_uint64(0)
        rc = self._L.shkz_submit(self.handle, ctypes.byref(grid), None if params is None else ctypes.byref(params),
                                 ctypes.byref(t))
        if rc != 0:
            raise ShkzError(rc, last_error())
        return int(t.value)

    def poll(self, ticket: int):
        """(status, Result): status PENDING while queued or running."""
        r = Result()
        rc = self._L.shkz_poll(self.handle, ticket, ctypes.byref(r))
        return rc, r

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        r = Result()
        rc = self._L.shkz_wait(self.handle, ticket, timeout_ms, ctypes.byref(r))
        return rc, r

    def cancel(self, ticket: int) -> int:
        """OK: cancelled while queued (the ticket then reports ERR_CANCELLED); PENDING: already started
        or done (its result stands); ERR_NOT_FOUND: unknown ticket."""
        return self._L.shkz_cancel(self.handle, ticket)

    def release_scene(self, scene_id: int):
        rc = self._L.shkz_server_release_scene(self.handle, scene_id)
        if rc != 0:
            raise ShkzError(rc, last_error())

    def stats(self) -> dict:
        st = ServerStats()
        rc = self._L.shkz_server_get_stats(self.handle, ctypes.byref(st))
        if rc != 0:
            raise ShkzError(rc, last_error())
        return st.as_dict()


class Context:
    """One shkz_ctx on a HIP device (the module instance of macpressuresolver3)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._L = lib()
        h = _vp()
        rc = self._L.shkz_create(device, hbm_budget_bytes, ctypes.byref(h))
        if rc != 0:
            raise ShkzError(rc, last_error())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self._L.shkz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ projection
    def project(self, scene: Scene, params: Params | None = None, pressure: bool = True, pressure64: bool = False,
                scene_id: int = 0, active=None, active_out=None, unknown_out=None, check: bool = True):
        """Projects scene's velocities in place (host arrays).  Returns (result, P32, P64).
        active_out: optional three uint8 face arrays that receive the post-projection activity;
        unknown_out: optional uint8 cell array that receives the unknown mask."""
        g = self._grid(scene, scene_id, active)
        if active_out is not None:
            for a in active_out:
                assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
            g.active_out_u, g.active_out_v, g.active_out_w = (_ptr(a) for a in active_out)
        if unknown_out is not None:
            assert unknown_out.dtype == np.uint8 and unknown_out.flags["C_CONTIGUOUS"]
            assert unknown_out.size == scene.w * scene.h * scene.d
            g.unknown_out = _ptr(unknown_out)
        P = np.zeros((scene.w, scene.h, scene.d), np.float32) if pressure else None
        P64 = np.zeros((scene.w, scene.h, scene.d), np.float64) if pressure64 else None
        g.pressure = _ptr(P)
        g.pressure64 = _ptr(P64)
        r = Result()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project(self.handle, ctypes.byref(g), ctypes.byref(prm), ctypes.byref(r))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return r, P, P64

    def project_batch(self, scenes, params: Params | None = None, pressure: bool = True, pressure64: bool = False,
                      scene_ids=None, check: bool = True):
        """Projects a batch of scenes in one call (host arrays, velocities updated in place).
        Returns (results, [P32], [P64])."""
        n = len(scenes)
        grids = (Grid * n)()
        Ps, P64s = [], []
        for i, sc in enumerate(scenes):
            grids[i] = self._grid(sc, 0 if scene_ids is None else scene_ids[i], None)
            P = np.zeros((sc.w, sc.h, sc.d), np.float32) if pressure else None
            P64 = np.zeros((sc.w, sc.h, sc.d), np.float64) if pressure64 else None
            grids[i].pressure = _ptr(P)
            grids[i].pressure64 = _ptr(P64)
            Ps.append(P)
            P64s.append(P64)
        res = (Result * n)()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project_batch(self.handle, grids, n, ctypes.byref(prm), res)
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return list(res), Ps, P64s

    @staticmethod
    def device_grid(w, h, d, dt, fluid, fluid_dtype, solid, solid_dtype, u, v, wv, pressure=None, pressure64=None,
                    scene_id=0, dx=0.0, active=None, active_out=None) -> Grid:
        """A Grid of device-resident buffers (ints or objects with data_ptr(), e.g. devmem.DeviceArray)."""
        g = Grid(w=w, h=h, d=d, dx=dx, dt=dt, fluid=_dp(fluid), fluid_dtype=fluid_dtype, solid_dtype=solid_dtype,
                 solid=_dp(solid), u=_dp(u), v=_dp(v), w_vel=_dp(wv), pressure=_dp(pressure),
                 pressure64=_dp(pressure64), scene_id=scene_id)
        if active is not None:
            g.active_u, g.active_v, g.active_w = (_dp(a) for a in active)
        if active_out is not None:
            g.active_out_u, g.active_out_v, g.active_out_w = (_dp(a) for a in active_out)
        return g

    def project_batch_device(self, grids, params: Params | None = None, stream=None, check: bool = True):
        """Batch projection of device-resident grids (a list of Grid from device_grid())."""
        n = len(grids)
        arr = (Grid * n)(*grids)
        res = (Result * n)()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project_batch_device(self.handle, arr, n, ctypes.byref(prm), res,
                                               None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return list(res)

    def project_device(self, w, h, d, dt, fluid, fluid_dtype, solid, solid_dtype, u, v, wv, params=None,
                       pressure=None, pressure64=None, scene_id=0, stream=None, dx=0.0, active=None, active_out=None,
                       check=True):
        """Projection on device-resident buffers (ints or objects with data_ptr(), e.g. devmem.DeviceArray)."""
        g = self.device_grid(w, h, d, dt, fluid, fluid_dtype, solid, solid_dtype, u, v, wv, pressure, pressure64,
                             scene_id, dx, active, active_out)
        r = Result()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project_device(self.handle, ctypes.byref(g), ctypes.byref(prm), ctypes.byref(r),
                                         None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return r

    # ------------------------------------------------------------ after the projection
    def extrapolate_velocity(self, w, h, d, vel, active, solid=None, width=1, constrain=True, dx=0.0,
                             check=True) -> int:
        """Velocity extrapolation (`width` passes) and, with `constrain`, the solid constraint on host
        arrays, in place: vel = (u, v, w) float32, active = three uint8 arrays, solid nodal or None."""
        for a in list(vel) + list(active):
            assert a.flags["C_CONTIGUOUS"]
        for a in vel:
            assert a.dtype == np.float32
        for a in active:
            assert a.dtype == np.uint8
        f = VelocityField(w=w, h=h, d=d, dx=dx, solid=_ptr(solid),
                          solid_dtype=F64 if (solid is not None and solid.dtype == np.float64) else F32,
                          u=_ptr(vel[0]), v=_ptr(vel[1]), w_vel=_ptr(vel[2]), active_u=_ptr(active[0]),
                          active_v=_ptr(active[1]), active_w=_ptr(active[2]))
        rc = self._L.shkz_extrapolate_velocity(self.handle, ctypes.byref(f), int(width), int(bool(constrain)))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    def extrapolate_velocity_device(self, w, h, d, vel, active, solid=None, solid_dtype=F32, width=1,
                                    constrain=True, dx=0.0, stream=None, check=True) -> int:
        """The same on device buffers (ints or objects with data_ptr()), ordered on `stream`."""
        f = VelocityField(w=w, h=h, d=d, dx=dx, solid=_dp(solid), solid_dtype=solid_dtype, u=_dp(vel[0]),
                          v=_dp(vel[1]), w_vel=_dp(vel[2]), active_u=_dp(active[0]), active_v=_dp(active[1]),
                          active_w=_dp(active[2]))
        rc = self._L.shkz_extrapolate_velocity_device(self.handle, ctypes.byref(f), int(width), int(bool(constrain)),
                                                      None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    @staticmethod
    def _advect_params(dt, fluid, fluid_dtype, maccormack, trim, background, device, weno=False):
        p = AdvectParams()
        lib().shkz_advect_default_params(ctypes.byref(p))
        p.dt = dt
        p.weno = int(bool(weno))
        p.maccormack = int(bool(maccormack))
        p.trim_narrowband = int(trim)
        p.background = background
        if fluid is not None:
            if device:
                p.fluid, p.fluid_dtype = _dp(fluid), fluid_dtype
            else:
                assert fluid.flags["C_CONTIGUOUS"]
                p.fluid, p.fluid_dtype = _ptr(fluid), F64 if fluid.dtype == np.float64 else F32
        return p

    @staticmethod
    def _vfield(w, h, d, vel, active, dx, device):
        conv = _dp if device else _ptr
        if not device:
            for a in list(vel) + list(active):
                assert a.flags["C_CONTIGUOUS"]
        return VelocityField(w=w, h=h, d=d, dx=dx, u=conv(vel[0]), v=conv(vel[1]), w_vel=conv(vel[2]),
                             active_u=conv(active[0]), active_v=conv(active[1]), active_w=conv(active[2]))

    def advect_velocity(self, w, h, d, vel, active, fluid, dt, maccormack=True, trim=1, dx=0.0, check=True,
                        weno=False) -> int:
        """macadvection3::advect_vector(u, u, fluid, dt) on host arrays, in place on vel."""
        f = self._vfield(w, h, d, vel, active, dx, False)
        p = self._advect_params(dt, fluid, None, maccormack, trim, 0.0, False, weno)
        rc = self._L.shkz_advect_velocity(self.handle, ctypes.byref(f), ctypes.byref(p))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    def advect_velocity_device(self, w, h, d, vel, active, fluid, fluid_dtype, dt, maccormack=True, trim=1, dx=0.0,
                               stream=None, check=True, weno=False) -> int:
        f = self._vfield(w, h, d, vel, active, dx, True)
        p = self._advect_params(dt, fluid, fluid_dtype, maccormack, trim, 0.0, True, weno)
        rc = self._L.shkz_advect_velocity_device(self.handle, ctypes.byref(f), ctypes.byref(p),
                                                 None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    def advect_scalar(self, w, h, d, vel, active, q, dt, active_q=None, fluid=None, maccormack=True, trim=1,
                      background=0.0, dx=0.0, check=True, weno=False) -> int:
        """macadvection3::advect_scalar(q, velocity, fluid, dt) on host arrays, in place on q (float32)."""
        assert q.dtype == np.float32 and q.flags["C_CONTIGUOUS"]
        f = self._vfield(w, h, d, vel, active, dx, False)
        p = self._advect_params(dt, fluid, None, maccormack, trim, background, False, weno)
        rc = self._L.shkz_advect_scalar(self.handle, ctypes.byref(f), _ptr(q), _ptr(active_q), ctypes.byref(p))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    def advect_scalar_device(self, w, h, d, vel, active, q, dt, active_q=None, fluid=None, fluid_dtype=F32,
                             maccormack=True, trim=1, background=0.0, dx=0.0, stream=None, check=True,
                             weno=False) -> int:
        f = self._vfield(w, h, d, vel, active, dx, True)
        p = self._advect_params(dt, fluid, fluid_dtype, maccormack, trim, background, True, weno)
        rc = self._L.shkz_advect_scalar_device(self.handle, ctypes.byref(f), _dp(q), _dp(active_q), ctypes.byref(p),
                                               None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return rc

    def _grid(self, scene: Scene, scene_id: int, active) -> Grid:
        for a in (scene.fluid, scene.u, scene.v, scene.wv):
            assert a.flags["C_CONTIGUOUS"]
        g = Grid(w=scene.w, h=scene.h, d=scene.d, dx=0.0, dt=scene.dt, fluid=_ptr(scene.fluid),
                 fluid_dtype=F64 if scene.fluid.dtype == np.float64 else F32,
                 solid_dtype=F64 if (scene.solid is not None and scene.solid.dtype == np.float64) else F32,
                 solid=_ptr(scene.solid), u=_ptr(scene.u), v=_ptr(scene.v), w_vel=_ptr(scene.wv), scene_id=scene_id)
        if active is not None:
            g.active_u, g.active_v, g.active_w = (_ptr(a) for a in active)
        return g

    # ------------------------------------------------------------------ 2D
    @staticmethod
    def grid2d(scene: Scene2D, scene_id: int = 0, active=None) -> Grid2D:
        for a in (scene.fluid, scene.u, scene.v):
            assert a.flags["C_CONTIGUOUS"]
        g = Grid2D(w=scene.w, h=scene.h, dx=0.0, dt=scene.dt, fluid=_ptr(scene.fluid),
                   fluid_dtype=F64 if scene.fluid.dtype == np.float64 else F32,
                   solid_dtype=F64 if (scene.solid is not None and scene.solid.dtype == np.float64) else F32,
                   solid=_ptr(scene.solid), u=_ptr(scene.u), v=_ptr(scene.v), scene_id=scene_id)
        if active is not None:
            g.active_u, g.active_v = (_ptr(a) for a in active)
        return g

    def project2d(self, scene: Scene2D, params: Params | None = None, pressure: bool = True,
                  pressure64: bool = False, scene_id: int = 0, active=None, check: bool = True):
        """Projects a 2D scene's velocities in place (host arrays).  Returns (result, P32, P64)."""
        g = self.grid2d(scene, scene_id, active)
        P = np.zeros((scene.w, scene.h), np.float32) if pressure else None
        P64 = np.zeros((scene.w, scene.h), np.float64) if pressure64 else None
        g.pressure = _ptr(P)
        g.pressure64 = _ptr(P64)
        r = Result()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project2d(self.handle, ctypes.byref(g), ctypes.byref(prm), ctypes.byref(r))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return r, P, P64

    def project2d_batch_device(self, grids, params: Params | None = None, stream=None, check: bool = True):
        """Batch of 2D grids whose pointers are device addresses (Grid2D)."""
        n = len(grids)
        arr = (Grid2D * n)(*grids)
        res = (Result * n)()
        prm = params if params is not None else default_params()
        rc = self._L.shkz_project2d_batch_device(self.handle, arr, n, ctypes.byref(prm), res,
                                                 None if stream is None else int(stream))
        if rc != 0 and check:
            raise ShkzError(rc, last_error())
        return list(res)

    def export_system2d(self, scene: Scene2D, params: Params | None = None) -> dict:
        """Assembled FP64 2D system on dense cell arrays: diag, gx, gy (low-face couplings), rhs."""
        g = self.grid2d(scene)
        shape = (scene.w, scene.h)
        out = {k: np.zeros(shape, np.float64) for k in ("diag", "gx", "gy", "rhs")}
        prm = params if params is not None else default_params()
        rc = self._L.shkz_export_system2d(self.handle, ctypes.byref(g), ctypes.byref(prm),
                                          *(out[k].ctypes.data for k in ("diag", "gx", "gy", "rhs")))
        if rc != 0:
            raise ShkzError(rc, last_error())
        return out

    def export_system(self, scene: Scene, params: Params | None = None) -> dict:
        """Assembled FP64 system on dense cell arrays: diag, gx, gy, gz (low-face couplings), rhs."""
        g = self._grid(scene, 0, None)
        shape = (scene.w, scene.h, scene.d)
        out = {k: np.zeros(shape, np.float64) for k in ("diag", "gx", "gy", "gz", "rhs")}
        prm = params if params is not None else default_params()
        rc = self._L.shkz_export_system(self.handle, ctypes.byref(g), ctypes.byref(prm), *(out[k].ctypes.data for k in
                                                                                         ("diag", "gx", "gy", "gz",
                                                                                          "rhs")))
        if rc != 0:
            raise ShkzError(rc, last_error())
        return out

    def release_scene(self, scene_id: int):
        rc = self._L.shkz_release_scene(self.handle, scene_id)
        if rc != 0:
            raise ShkzError(rc, last_error())

    def device_bytes(self) -> int:
        return int(self._L.shkz_device_bytes(self.handle))

    def inject_alloc_failure(self, after: int):
        """Test hook: the device allocation after `after` more successful ones fails (after < 0: off)."""
        rc = self._L.shkz_inject_alloc_failure(self.handle, after)
        if rc != 0:
            raise ShkzError(rc, last_error())

    # ------------------------------------------------------------------ kernel timers
    def set_kernel_timing(self, on: bool):
        self._L.shkz_set_kernel_timing(self.handle, 1 if on else 0)

    def reset_kernel_stats(self):
        self._L.shkz_reset_kernel_stats(self.handle)

    def kernel_stats(self) -> dict:
        n = ctypes.c_int(0)
        buf = (KernelStat * 128)()
        rc = self._L.shkz_get_kernel_stats(self.handle, buf, 128, ctypes.byref(n))
        if rc != 0:
            raise ShkzError(rc, last_error())
        out = {}
        for i in range(min(n.value, 128)):
            s = buf[i]
            out[s.name.decode()] = {"launches": int(s.launches), "total_ms": float(s.total_ms), "bytes": float(s.bytes)}
        return out
