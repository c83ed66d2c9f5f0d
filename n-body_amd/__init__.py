"""Python mirror of the reference's Solver API over the nbg C ABI (libnbg.so).

The reference (JacksonCampolattaro/n-body) is C++: a ``Simulation`` holds the
bodies, a ``Gravity`` rule holds G and epsilon, and a tree solver such as
``LinearBVHSolver(simulation, rule)`` exposes ``step()``,
``updateAccelerations()``, ``theta()``, ``maxLeafSize()`` and the static
``timeStep()`` (src/NBody/Simulation/Solver.h:26-187,
src/NBody/Simulation/Solvers/ActiveTreeSolver.h:16-82).  This module keeps
those names so tests and tools read like the reference's own code; every
call goes to the HIP kernels through ``include/nbg.h``.  There is no CPU
fallback: if libnbg.so is missing or no GPU is present, calls raise.
"""

from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbg.so")

BARNES_HUT = 0
LINEAR_BVH = 1
NAIVE = 2
BARNES_HUT_QUADRUPOLE = 3
LINEAR_BVH_QUADRUPOLE = 4
IMPLICIT_FMM = 5
SUM_ORDERED = 0
SUM_FAST = 1

PHASES = ("bbox", "keys", "sort", "topology", "gather", "summaries", "walk", "update")

# Every symbol include/nbg.h declares.
EXPORTS = (
    "nbg_params_default", "nbg_last_error", "nbg_version", "nbg_device_count",
    "nbg_ctx_create", "nbg_ctx_destroy", "nbg_sys_create", "nbg_sys_destroy",
    "nbg_sys_set_params", "nbg_sys_get_params", "nbg_sys_set_state", "nbg_sys_get_state",
    "nbg_sys_update_accelerations", "nbg_sys_get_accelerations", "nbg_sys_accelerations_at", "nbg_sys_accelerations_at_noalloc", "nbg_sys_step",
    "nbg_sys_synchronize", "nbg_sys_counts", "nbg_sys_counts4", "nbg_sys_enable_counts", "nbg_sys_enable_walk_stats",
    "nbg_sys_walk_stats", "nbg_sys_set_walk_split",
    "nbg_sys_enable_timing", "nbg_sys_phase_times",
    "nbg_sys_debug_bbox", "nbg_sys_debug_order", "nbg_sys_debug_tree", "nbg_sys_debug_quads", "nbg_direct_f64",
    "nbg_sys_potential_f64", "nbg_generate_plummer", "nbg_generate_cube",
    "nbg_batch_bytes", "nbg_batch_create", "nbg_batch_destroy", "nbg_batch_add", "nbg_batch_remove",
    "nbg_batch_set_params", "nbg_batch_info", "nbg_batch_get", "nbg_batch_update_accelerations", "nbg_batch_step",
    "nbg_batch_synchronize", "nbg_batch_counts", "nbg_batch_enable_counts", "nbg_batch_use_graph",
    "nbg_batch_autotune", "nbg_batch_autotune_grid", "nbg_batch_set_targets_per_lane", "nbg_batch_set_walk_split", "nbg_batch_nonfinite",
    "nbg_scenario_load_json", "nbg_scenario_load_tipsy", "nbg_scenario_count", "nbg_scenario_get", "nbg_scenario_free",
    "nbg_server_create", "nbg_server_create_express", "nbg_server_destroy", "nbg_server_capacity", "nbg_server_submit", "nbg_server_wait",
    "nbg_server_stats", "nbg_server_checkpoint", "nbg_server_suspend", "nbg_server_resume",
)


class NbgError(RuntimeError):
    """A negative status from the C ABI, with nbg_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"nbg error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """nbg_params (include/nbg.h)."""

    _fields_ = [
        ("solver", ctypes.c_int32),
        ("theta", ctypes.c_float),
        ("leaf_size", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("morton_bits", ctypes.c_int32),
        ("G", ctypes.c_float),
        ("eps2", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("sum_mode", ctypes.c_int32),
        ("targets_per_lane", ctypes.c_int32),
    ]


_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    """Loads the in-tree libnbg.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "nbg_params_default": (None, [ctypes.POINTER(Params)]),
            "nbg_last_error": (ctypes.c_char_p, []),
            "nbg_version": (ctypes.c_char_p, []),
            "nbg_device_count": (ctypes.c_int, []),
            "nbg_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
            "nbg_ctx_destroy": (ctypes.c_int, [_vp]),
            "nbg_sys_create": (ctypes.c_int, [_vp, ctypes.POINTER(Params), ctypes.c_int64, ctypes.POINTER(_vp)]),
            "nbg_sys_destroy": (ctypes.c_int, [_vp]),
            "nbg_sys_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(Params)]),
            "nbg_sys_get_params": (ctypes.c_int, [_vp, ctypes.POINTER(Params)]),
            "nbg_sys_set_state": (ctypes.c_int, [_vp] + [_f32p] * 7),
            "nbg_sys_get_state": (ctypes.c_int, [_vp] + [_f32p] * 6),
            "nbg_sys_update_accelerations": (ctypes.c_int, [_vp] + [_f32p] * 3),
            "nbg_sys_get_accelerations": (ctypes.c_int, [_vp] + [_f32p] * 3),
            "nbg_sys_accelerations_at": (ctypes.c_int, [_vp, ctypes.c_int64] + [_f32p] * 6),
            "nbg_sys_accelerations_at_noalloc": (ctypes.c_int, [_vp, ctypes.c_int64] + [_f32p] * 6),
            "nbg_batch_bytes": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int32]),
            "nbg_batch_create": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.POINTER(_vp)]),
            "nbg_batch_destroy": (ctypes.c_int, [_vp]),
            "nbg_batch_add": (ctypes.c_int, [_vp, ctypes.POINTER(Params), ctypes.c_int64] + [_f32p] * 7
                              + [ctypes.POINTER(ctypes.c_int32)]),
            "nbg_batch_remove": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_batch_set_params": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(Params)]),
            "nbg_batch_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
            "nbg_batch_get": (ctypes.c_int, [_vp, ctypes.c_int32] + [_f32p] * 9),
            "nbg_batch_update_accelerations": (ctypes.c_int, [_vp]),
            "nbg_batch_nonfinite": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_int32)]),
            "nbg_batch_step": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_batch_synchronize": (ctypes.c_int, [_vp]),
            "nbg_batch_counts": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
            "nbg_batch_enable_counts": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_batch_use_graph": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_batch_autotune": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.POINTER(ctypes.c_int32), _f32p]),
            "nbg_batch_autotune_grid": (ctypes.c_int, [_vp, _i32p, ctypes.c_int32, _i32p, ctypes.c_int32,
                                                       ctypes.c_int32, _i32p, _i32p, _f32p]),
            "nbg_batch_set_targets_per_lane": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_batch_set_walk_split": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_scenario_load_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
            "nbg_scenario_load_tipsy": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
            "nbg_scenario_count": (ctypes.c_int64, [_vp]),
            "nbg_scenario_get": (ctypes.c_int, [_vp] + [_f32p] * 7 + [ctypes.POINTER(ctypes.c_uint8)]),
            "nbg_scenario_free": (ctypes.c_int, [_vp]),
            "nbg_server_create": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.POINTER(_vp)]),
            "nbg_server_create_express": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                         ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                                         ctypes.POINTER(_vp)]),
            "nbg_server_destroy": (ctypes.c_int, [_vp]),
            "nbg_server_capacity": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64),
                                                   ctypes.POINTER(ctypes.c_int32)]),
            "nbg_server_submit": (ctypes.c_int, [_vp, ctypes.POINTER(Params), ctypes.c_int64] + [_f32p] * 7
                                  + [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
            "nbg_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double] + [_f32p] * 6
                                + [ctypes.POINTER(ctypes.c_double)]),
            "nbg_server_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
            "nbg_server_checkpoint": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                                     ctypes.POINTER(ctypes.c_int64)]),
            "nbg_server_suspend": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                                  ctypes.POINTER(ctypes.c_int64)]),
            "nbg_server_resume": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.POINTER(ctypes.c_int64)]),
            "nbg_sys_step": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_sys_synchronize": (ctypes.c_int, [_vp]),
            "nbg_sys_counts": (ctypes.c_int, [_vp, _u64p, _u64p]),
            "nbg_sys_counts4": (ctypes.c_int, [_vp, _u64p]),
            "nbg_sys_walk_stats": (ctypes.c_int, [_vp, _u64p]),
            "nbg_sys_set_walk_split": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_sys_enable_walk_stats": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_sys_enable_counts": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_sys_enable_timing": (ctypes.c_int, [_vp, ctypes.c_int32]),
            "nbg_sys_phase_times": (ctypes.c_int, [_vp, _f32p, ctypes.c_int32]),
            "nbg_sys_debug_bbox": (ctypes.c_int, [_vp, _f32p]),
            "nbg_sys_debug_order": (ctypes.c_int, [_vp, _u64p, _u32p]),
            "nbg_sys_debug_tree": (ctypes.c_int, [_vp, _i64p, _i64p, _i64p, _i32p, _i32p, _f32p, _f32p]),
            "nbg_sys_debug_quads": (ctypes.c_int, [_vp, ctypes.c_int64, _f32p]),
            "nbg_direct_f64": (ctypes.c_int, [_vp, ctypes.c_int64] + [_f32p] * 4 + [ctypes.c_int64] + [_f32p] * 3
                               + [ctypes.c_double, ctypes.c_double] + [_f64p] * 3),
            "nbg_sys_potential_f64": (ctypes.c_int, [_vp, _f64p]),
            "nbg_generate_plummer": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double] + [_f32p] * 7),
            "nbg_generate_cube": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double]
                                  + [_f32p] * 7),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> int:
    if rc < 0:
        raise NbgError(rc, lib().nbg_last_error().decode())
    return rc


def _ptr(a, t=_f32p):
    return None if a is None else a.ctypes.data_as(t)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def default_params() -> Params:
    p = Params()
    lib().nbg_params_default(ctypes.byref(p))
    return p


def device_count() -> int:
    return lib().nbg_device_count()


# ------------------------------------------------------------------ generators

@dataclass
class Bodies:
    """SoA FP32 body state: the Simulation's Position / Velocity / Mass."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    m: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "Bodies":
        return Bodies(*(a.copy() for a in (self.x, self.y, self.z, self.m, self.vx, self.vy, self.vz)))

    @staticmethod
    def empty(n: int) -> "Bodies":
        return Bodies(*(np.zeros(n, np.float32) for _ in range(7)))


def plummer(n: int, seed: int, a: float = 1.0, M: float = 1.0, G: float = 1.0, rmax: float = 10.0) -> Bodies:
    """Plummer sphere (nbg_generate_plummer)."""
    b = Bodies.empty(n)
    _check(lib().nbg_generate_plummer(seed, n, a, M, G, rmax, *[_ptr(v) for v in
                                                                  (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)]))
    return b


def cube(n: int, seed: int, side: float = 2.0, M: float = 1.0) -> Bodies:
    """Uniform cube at rest (nbg_generate_cube)."""
    b = Bodies.empty(n)
    _check(lib().nbg_generate_cube(seed, n, side, M, *[_ptr(v) for v in (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)]))
    return b


# ------------------------------------------------------------------ C ABI objects

class Context:
    """nbg_ctx: one GPU."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().nbg_ctx_create(device, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().nbg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def direct_f64(self, src: Bodies, tx, ty, tz, G: float, eps2: float):
        """FP64 direct-sum accelerations at the target points (nbg_direct_f64)."""
        tx, ty, tz = _f32(tx), _f32(ty), _f32(tz)
        nt = tx.shape[0]
        out = [np.zeros(nt, np.float64) for _ in range(3)]
        _check(lib().nbg_direct_f64(self.handle, src.n, _ptr(src.x), _ptr(src.y), _ptr(src.z), _ptr(src.m), nt,
                                    _ptr(tx), _ptr(ty), _ptr(tz), G, eps2, *[_ptr(o, _f64p) for o in out]))
        return np.stack(out, axis=1)


class System:
    """nbg_sys: one set of bodies plus one tree solver's device state."""

    def __init__(self, ctx: Context, params: Params, n: int):
        h = _vp()
        self._ctx = ctx
        _check(lib().nbg_sys_create(ctx.handle, ctypes.byref(params), n, ctypes.byref(h)))
        self.handle = h
        self.n = n

    def close(self):
        if self.handle:
            lib().nbg_sys_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> Params:
        p = Params()
        _check(lib().nbg_sys_get_params(self.handle, ctypes.byref(p)))
        return p

    @params.setter
    def params(self, p: Params):
        _check(lib().nbg_sys_set_params(self.handle, ctypes.byref(p)))

    def set_state(self, b: Bodies):
        arrs = [_f32(v) for v in (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)]
        _check(lib().nbg_sys_set_state(self.handle, *[_ptr(a) for a in arrs]))

    def get_state(self) -> Bodies:
        out = Bodies.empty(self.n)
        _check(lib().nbg_sys_get_state(self.handle, *[_ptr(v) for v in (out.x, out.y, out.z, out.vx, out.vy, out.vz)]))
        return out

    def update_accelerations(self) -> np.ndarray:
        a = [np.zeros(self.n, np.float32) for _ in range(3)]
        _check(lib().nbg_sys_update_accelerations(self.handle, *[_ptr(v) for v in a]))
        return np.stack(a, axis=1)

    def accelerations(self) -> np.ndarray:
        a = [np.zeros(self.n, np.float32) for _ in range(3)]
        _check(lib().nbg_sys_get_accelerations(self.handle, *[_ptr(v) for v in a]))
        return np.stack(a, axis=1)

    def accelerations_at(self, tx, ty, tz) -> np.ndarray:
        """Accelerations at passive points (bodies without Mass) from the
        tree of the current state."""
        t = [_f32(v) for v in (tx, ty, tz)]
        nt = len(t[0])
        a = [np.zeros(nt, np.float32) for _ in range(3)]
        _check(lib().nbg_sys_accelerations_at(self.handle, nt, *[_ptr(v) for v in t], *[_ptr(v) for v in a]))
        return np.stack(a, axis=1)

    def accelerations_at_noalloc(self, tx, ty, tz) -> np.ndarray:
        """accelerations_at through the allocation-free entry point (static
        device scratch, at most NOALLOC_MAX_POINTS targets)."""
        t = [_f32(v) for v in (tx, ty, tz)]
        nt = len(t[0])
        a = [np.zeros(nt, np.float32) for _ in range(3)]
        _check(lib().nbg_sys_accelerations_at_noalloc(self.handle, nt, *[_ptr(v) for v in t],
                                                      *[_ptr(v) for v in a]))
        return np.stack(a, axis=1)

    def step(self, nsteps: int = 1):
        _check(lib().nbg_sys_step(self.handle, nsteps))

    def synchronize(self):
        _check(lib().nbg_sys_synchronize(self.handle))

    def counts(self):
        pp, np_ = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().nbg_sys_counts(self.handle, ctypes.byref(pp), ctypes.byref(np_)))
        return int(pp.value), int(np_.value)

    def counts4(self):
        """(pp, np, pn, nn) as SimpleTrackingRule counts them."""
        out = (ctypes.c_uint64 * 4)()
        _check(lib().nbg_sys_counts4(self.handle, out))
        return tuple(int(v) for v in out)

    def enable_counts(self, on: bool = True):
        _check(lib().nbg_sys_enable_counts(self.handle, int(on)))

    def enable_walk_stats(self, on: bool = True):
        _check(lib().nbg_sys_enable_walk_stats(self.handle, int(on)))

    def set_walk_split(self, waves: int = 0):
        """Waves per group of 64 targets in the FAST linear-BVH walk (0:
        chosen by size; results are the same bits for every choice)."""
        _check(lib().nbg_sys_set_walk_split(self.handle, int(waves)))

    def walk_stats(self) -> dict:
        out = np.zeros(4, np.uint64)
        _check(lib().nbg_sys_walk_stats(self.handle, _ptr(out, _u64p)))
        v = [int(x) for x in out]
        if self.params.solver == IMPLICIT_FMM:
            return {"pool_peak": v[0], "frontier_peak": v[1], "steps": v[2], "pool_capacity": v[3]}
        return {"node_visits": v[0], "active_lanes": v[1], "leaf_visits": v[2], "leaf_lanes": v[3]}

    def enable_timing(self, on: bool = True):
        _check(lib().nbg_sys_enable_timing(self.handle, int(on)))

    def phase_times(self) -> dict:
        ms = np.zeros(len(PHASES), np.float32)
        k = _check(lib().nbg_sys_phase_times(self.handle, _ptr(ms), len(PHASES)))
        return {PHASES[i]: float(ms[i]) for i in range(k)}

    def debug_bbox(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        _check(lib().nbg_sys_debug_bbox(self.handle, _ptr(out)))
        return out

    def debug_order(self):
        keys = np.zeros(self.n, np.uint64)
        perm = np.zeros(self.n, np.uint32)
        _check(lib().nbg_sys_debug_order(self.handle, _ptr(keys, _u64p), _ptr(perm, _u32p)))
        return keys, perm

    def debug_tree(self) -> dict:
        nn = ctypes.c_int64()
        _check(lib().nbg_sys_debug_tree(self.handle, ctypes.byref(nn), None, None, None, None, None, None))
        k = nn.value
        first = np.zeros(k, np.int64)
        count = np.zeros(k, np.int64)
        leaf = np.zeros(k, np.int32)
        parent = np.zeros(k, np.int32)
        com4 = np.zeros((k, 4), np.float32)
        geo6 = np.zeros((k, 6), np.float32)
        _check(lib().nbg_sys_debug_tree(self.handle, ctypes.byref(nn), _ptr(first, _i64p), _ptr(count, _i64p),
                                        _ptr(leaf, _i32p), _ptr(parent, _i32p), _ptr(com4), _ptr(geo6)))
        return dict(first=first, count=count, leaf=leaf.astype(bool), parent=parent, com4=com4, geo6=geo6)

    def debug_quads(self) -> np.ndarray:
        """Per-node second moments (xx, xy, xz, yy, yz, zz) in debug_tree's order."""
        nn = ctypes.c_int64()
        _check(lib().nbg_sys_debug_tree(self.handle, ctypes.byref(nn), None, None, None, None, None, None))
        q = np.zeros((nn.value, 6), np.float32)
        _check(lib().nbg_sys_debug_quads(self.handle, nn.value, _ptr(q)))
        return q

    def potential_f64(self) -> float:
        w = ctypes.c_double()
        _check(lib().nbg_sys_potential_f64(self.handle, ctypes.byref(w)))
        return float(w.value)


# ------------------------------------------------------------------ batched forest

class Batch:
    """nbg_batch: independent linear-BVH systems (all monopole or all
    quadrupole) stepped together in one segmented build and one walk launch
    per step."""

    def __init__(self, ctx: Context, max_bodies: int, max_systems: int, sum_mode: int = SUM_FAST):
        h = _vp()
        _check(lib().nbg_batch_create(ctx.handle, max_bodies, max_systems, sum_mode, ctypes.byref(h)))
        self.handle = h
        self._n = {}

    @staticmethod
    def bytes_for(max_bodies: int, max_systems: int) -> int:
        return int(lib().nbg_batch_bytes(max_bodies, max_systems))

    def close(self):
        if self.handle:
            lib().nbg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, params: Params, b: Bodies) -> int:
        arrs = [_f32(v) for v in (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)]
        sid = ctypes.c_int32()
        _check(lib().nbg_batch_add(self.handle, ctypes.byref(params), b.n, *[_ptr(a) for a in arrs],
                                   ctypes.byref(sid)))
        self._n[sid.value] = b.n
        return sid.value

    def remove(self, sid: int):
        _check(lib().nbg_batch_remove(self.handle, sid))
        del self._n[sid]

    def set_params(self, sid: int, params: Params):
        _check(lib().nbg_batch_set_params(self.handle, sid, ctypes.byref(params)))

    def info(self):
        ns, nb = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().nbg_batch_info(self.handle, ctypes.byref(ns), ctypes.byref(nb)))
        return ns.value, nb.value

    def get(self, sid: int, accelerations: bool = False):
        n = self._n[sid]
        out = Bodies.empty(n)
        acc = [np.zeros(n, np.float32) for _ in range(3)] if accelerations else [None] * 3
        _check(lib().nbg_batch_get(self.handle, sid, *[_ptr(v) for v in (out.x, out.y, out.z, out.vx, out.vy, out.vz)],
                                   *[_ptr(v) if v is not None else None for v in acc]))
        return (out, np.stack(acc, axis=1)) if accelerations else out

    def nonfinite(self):
        """Ids of the systems whose positions or velocities hold NaN or inf."""
        cap = max(1, len(self._n))
        ids = np.zeros(cap, np.int32)
        k = ctypes.c_int32()
        _check(lib().nbg_batch_nonfinite(self.handle, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                         ctypes.byref(k)))
        return [int(v) for v in ids[:min(k.value, cap)]]

    def update_accelerations(self):
        _check(lib().nbg_batch_update_accelerations(self.handle))

    def step(self, nsteps: int = 1):
        _check(lib().nbg_batch_step(self.handle, nsteps))

    def synchronize(self):
        _check(lib().nbg_batch_synchronize(self.handle))

    def counts(self):
        pp, np_ = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().nbg_batch_counts(self.handle, ctypes.byref(pp), ctypes.byref(np_)))
        return int(pp.value), int(np_.value)

    def enable_counts(self, on: bool = True):
        _check(lib().nbg_batch_enable_counts(self.handle, int(on)))

    def use_graph(self, on: bool = True):
        _check(lib().nbg_batch_use_graph(self.handle, int(on)))

    def autotune_grid(self, leaf_sizes=(8, 16, 32, 64), tpls=(1, 2, 4), repeats: int = 2):
        """Leaf size x targets per lane autotuner; returns (best leaf, best
        targets per lane, {(leaf, tpl): ms}) with untimed pairs left out."""
        ls = np.asarray(leaf_sizes, np.int32)
        ts = np.asarray(tpls, np.int32)
        ms = np.zeros(len(ls) * len(ts), np.float32)
        bl, bt = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().nbg_batch_autotune_grid(self.handle, ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ls),
                                             ts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ts), repeats,
                                             ctypes.byref(bl), ctypes.byref(bt), _ptr(ms)))
        table = {(int(l), int(t)): float(ms[i * len(ts) + j]) for i, l in enumerate(ls) for j, t in enumerate(ts)
                 if ms[i * len(ts) + j] >= 0}
        return bl.value, bt.value, table

    def set_targets_per_lane(self, tpl: int):
        _check(lib().nbg_batch_set_targets_per_lane(self.handle, int(tpl)))

    def set_walk_split(self, waves: int = 0):
        """Waves per group of 64 targets in the FAST walks: 0 by size (forests
        of at most 65,536 bodies, at least half of them in systems of 4,096 or
        more, split), 1 off, or 2, 4, 8, 16."""
        _check(lib().nbg_batch_set_walk_split(self.handle, int(waves)))

    def autotune(self, leaf_sizes=(8, 16, 32, 64), repeats: int = 2):
        ls = np.asarray(leaf_sizes, np.int32)
        ms = np.zeros(len(ls), np.float32)
        best = ctypes.c_int32()
        _check(lib().nbg_batch_autotune(self.handle, ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(ls),
                                        repeats, ctypes.byref(best), _ptr(ms)))
        return best.value, {int(a): float(t) for a, t in zip(ls, ms)}


SCENARIO_POSITION, SCENARIO_VELOCITY, SCENARIO_ACTIVE, SCENARIO_PASSIVE = 1, 2, 4, 8


def load_scenario(path: str, kind: str = "json"):
    """A reference scenario file read by the library's C reader
    (nbg_scenario_load_json / _tipsy): (x, y, z, vx, vy, vz, m, flags)
    arrays in file order."""
    h = _vp()
    fn = lib().nbg_scenario_load_json if kind == "json" else lib().nbg_scenario_load_tipsy
    _check(fn(os.fsencode(path), ctypes.byref(h)))
    try:
        n = lib().nbg_scenario_count(h)
        arrs = [np.zeros(n, np.float32) for _ in range(7)]
        flags = np.zeros(n, np.uint8)
        _check(lib().nbg_scenario_get(h, *[_ptr(a) for a in arrs], flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    finally:
        lib().nbg_scenario_free(h)
    return (*arrs, flags)


class Server:
    """nbg_server: continuous batching of submitted jobs over one forest,
    within an HBM byte budget; with express_max_bodies > 0, jobs up to that
    size run in a second forest (express_bytes of the budget,
    express_systems slots) that steps concurrently with the first
    (nbg_server_create_express)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, max_systems: int = 4096, sum_mode: int = SUM_FAST,
                 express_max_bodies: int = 0, express_bytes: int = 0, express_systems: int = 0):
        h = _vp()
        if express_max_bodies > 0:
            _check(lib().nbg_server_create_express(ctx.handle, int(hbm_budget_bytes), max_systems, sum_mode,
                                                   int(express_max_bodies), int(express_bytes), express_systems,
                                                   ctypes.byref(h)))
        else:
            _check(lib().nbg_server_create(ctx.handle, int(hbm_budget_bytes), max_systems, sum_mode,
                                           ctypes.byref(h)))
        self.handle = h
        self._n = {}

    def close(self):
        if self.handle:
            lib().nbg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def capacity(self):
        nb, ns = ctypes.c_int64(), ctypes.c_int32()
        _check(lib().nbg_server_capacity(self.handle, ctypes.byref(nb), ctypes.byref(ns)))
        return nb.value, ns.value

    def submit(self, params: Params, b: Bodies, nsteps: int) -> int:
        arrs = [_f32(v) for v in (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)]
        job = ctypes.c_int64()
        _check(lib().nbg_server_submit(self.handle, ctypes.byref(params), b.n, *[_ptr(a) for a in arrs], nsteps,
                                       ctypes.byref(job)))
        self._n[job.value] = b.n
        return job.value

    def wait(self, job: int, timeout: float = -1.0):
        """(final Bodies without masses, {queued, device, latency} seconds)."""
        n = self._n[job]
        out = Bodies.empty(n)
        t = (ctypes.c_double * 3)()
        _check(lib().nbg_server_wait(self.handle, job, float(timeout),
                                     *[_ptr(v) for v in (out.x, out.y, out.z, out.vx, out.vy, out.vz)], t))
        del self._n[job]
        return out, {"queued": t[0], "device": t[1], "latency": t[2]}

    def stats(self) -> dict:
        o = (ctypes.c_double * 8)()
        _check(lib().nbg_server_stats(self.handle, o))
        keys = ("jobs_done", "steps_run", "body_steps", "busy_s", "wall_s", "peak_bodies", "peak_systems", "waiting")
        return {k: float(v) for k, v in zip(keys, o)}

    def _snapshot(self, fn, job: int) -> bytes:
        size = ctypes.c_int64()
        _check(fn(self.handle, job, None, 0, ctypes.byref(size)))
        buf = np.empty(size.value, dtype=np.uint8)
        _check(fn(self.handle, job, buf.ctypes.data_as(ctypes.c_void_p), size.value, ctypes.byref(size)))
        return buf.tobytes()

    def checkpoint(self, job: int) -> bytes:
        """The job's state at its forest's next step boundary, as a snapshot
        (nbg_server_checkpoint); the job runs on."""
        return self._snapshot(lib().nbg_server_checkpoint, job)

    def suspend(self, job: int) -> bytes:
        """The job's snapshot, and the job leaves the server
        (nbg_server_suspend)."""
        snap = self._snapshot(lib().nbg_server_suspend, job)
        self._n.pop(job, None)
        return snap

    def resume(self, snapshot: bytes) -> int:
        """Queues the job a snapshot holds for its remaining steps; returns
        its new id (nbg_server_resume)."""
        buf = np.frombuffer(snapshot, dtype=np.uint8).copy()  # aligned
        job = ctypes.c_int64()
        _check(lib().nbg_server_resume(self.handle, buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                       ctypes.byref(job)))
        self._n[job.value] = snapshot_info(snapshot)["n"]
        return job.value


def snapshot_info(snapshot: bytes) -> dict:
    """Header fields of a job snapshot (nbg_server_checkpoint's format)."""
    if len(snapshot) < 64 or snapshot[:8] != b"NBGSNAP1":
        raise ValueError("not a job snapshot")
    version, params_bytes, n, total, done = struct.unpack_from("<IIqii", snapshot, 8)
    return {"version": version, "params_bytes": params_bytes, "n": n, "steps_total": total, "steps_done": done}


# ------------------------------------------------------------------ reference-shaped facade

class Gravity:
    """Gravity(G, epsilon) rule (src/NBody/Physics/Rules/Gravity.h:185-195).

    As in the reference, ``epsilon`` is added to r^2 (it is a squared
    softening length)."""

    def __init__(self, G: float = 1.0, epsilon: float = 0.00001):
        self.G = float(G)
        self.epsilon = float(epsilon)

    def g(self) -> float:
        return self.G


class Simulation:
    """The bodies a solver reads and writes (src/NBody/Simulation/Simulation.h:45-227),
    held as SoA FP32 arrays."""

    def __init__(self, bodies: Bodies):
        self.bodies = bodies
        self.acceleration = np.zeros((bodies.n, 3), np.float32)

    def particleCount(self) -> int:
        return self.bodies.n


class _TreeSolver:
    _kind = LINEAR_BVH
    _dt = 0.001  # Solver::_dt is static and shared by all solvers (Solver.h:34)

    def __init__(self, simulation: Simulation, rule: Gravity, context: Context | None = None):
        self._simulation = simulation
        self._rule = rule
        self._ctx = context or Context(0)
        self._params = default_params()
        self._params.solver = self._kind
        self._sys: System | None = None
        self._dirty = True

    # -- reference accessors
    @staticmethod
    def timeStep() -> float:
        return _TreeSolver._dt

    @staticmethod
    def setTimeStep(dt: float):
        _TreeSolver._dt = float(dt)

    def theta(self) -> float:
        return self._params.theta

    def setTheta(self, theta: float):
        self._params.theta = theta

    def maxLeafSize(self) -> int:
        return self._params.leaf_size

    def setMaxLeafSize(self, n: int):
        self._params.leaf_size = n

    def setSumMode(self, mode: int):
        self._params.sum_mode = mode

    def setMortonBits(self, bits: int):
        self._params.morton_bits = bits

    def simulation(self) -> Simulation:
        return self._simulation

    def rule(self) -> Gravity:
        return self._rule

    # -- device plumbing
    def _system(self) -> System:
        p = self._params
        p.G = self._rule.G
        p.eps2 = self._rule.epsilon
        p.dt = _TreeSolver._dt
        n = self._simulation.bodies.n
        if self._sys is None or self._sys.n != n:
            self._sys = System(self._ctx, p, n)
            self._dirty = True
        else:
            self._sys.params = p
        if self._dirty:
            self._sys.set_state(self._simulation.bodies)
            self._dirty = False
        return self._sys

    def updateAccelerations(self):
        """ActiveTreeSolver::updateAccelerations (ActiveTreeSolver.h:33-72)."""
        self._simulation.acceleration = self._system().update_accelerations()

    def step(self):
        """Solver::step (Solver.h:103-113): accelerations, velocities, positions."""
        s = self._system()
        s.step(1)
        st = s.get_state()
        b = self._simulation.bodies
        b.x, b.y, b.z, b.vx, b.vy, b.vz = st.x, st.y, st.z, st.vx, st.vy, st.vz
        self._simulation.acceleration = s.accelerations()

    def counts(self):
        return self._system().counts() if self._sys is not None else (0, 0)


class LinearBVHSolver(_TreeSolver):
    """LinearBVHSolver (src/NBody/Simulation/Solvers/LinearBVHSolver.h:18-28)."""

    _kind = LINEAR_BVH

    def id(self) -> str:
        return "gpu-linear-bvh"

    def name(self) -> str:
        return "Linear-BVH (MI355X)"


class BarnesHutSolver(_TreeSolver):
    """BarnesHutSolver (src/NBody/Simulation/Solvers/BarnesHutSolver.h:14-32)."""

    _kind = BARNES_HUT

    def id(self) -> str:
        return "gpu-barnes-hut"

    def name(self) -> str:
        return "Barnes-Hut (MI355X)"

    def maxDepth(self) -> int:
        return self._params.max_depth

    def setMaxDepth(self, d: int):
        self._params.max_depth = d


class QuadrupoleLinearBVHSolver(_TreeSolver):
    """QuadrupoleLinearBVHSolver = MultipoleLinearBVHSolver<2>
    (src/NBody/Simulation/Solvers/LinearBVHSolver.h:30-53)."""

    _kind = LINEAR_BVH_QUADRUPOLE

    def id(self) -> str:
        return "gpu-linear-bvh-4p"

    def name(self) -> str:
        return "Linear BVH (Quadrupole, MI355X)"


class QuadrupoleBarnesHutSolver(BarnesHutSolver):
    """QuadrupoleBarnesHutSolver = MultipoleBarnesHutSolver<2>
    (src/NBody/Simulation/Solvers/BarnesHutSolver.h:34-66)."""

    _kind = BARNES_HUT_QUADRUPOLE

    def id(self) -> str:
        return "gpu-barnes-hut-4p"

    def name(self) -> str:
        return "Barnes-Hut (Quadrupole, MI355X)"


class NaiveSolver(_TreeSolver):
    """NaiveSolver (src/NBody/Simulation/Solvers/NaiveSolver.h:14-52): direct
    summation over every body, no tree."""

    _kind = NAIVE

    def id(self) -> str:
        return "gpu-naive"

    def name(self) -> str:
        return "Naive (MI355X)"


class ImplicitFMMSolver(_TreeSolver):
    """ImplicitFMMSolver = MultipoleImplicitFMMSolver<1>
    (src/NBody/Simulation/Solvers/ImplicitFMMSolver.h:14-36): the implicit
    dual traversal (ImplicitDualTraversalSolver.h:17-70) over the Barnes-Hut
    octree with SideLengthOverDistance at theta 0.4.  FAST sum mode only."""

    _kind = IMPLICIT_FMM

    def __init__(self, simulation: Simulation, rule: Gravity, context: Context | None = None):
        super().__init__(simulation, rule, context)
        self._params.theta = 0.4  # ImplicitDualTraversalSolver's criterion (ImplicitDualTraversalSolver.h:22)
        self._params.sum_mode = SUM_FAST

    def id(self) -> str:
        return "gpu-implicit-fmm-2p"

    def name(self) -> str:
        return "Implicit Fast Multipole Method (1-pole, MI355X)"

    def maxDepth(self) -> int:
        return self._params.max_depth

    def setMaxDepth(self, d: int):
        self._params.max_depth = d

    def counts4(self):
        return self._system().counts4() if self._sys is not None else (0, 0, 0, 0)


SOLVERS = {"gpu-linear-bvh": LinearBVHSolver, "gpu-barnes-hut": BarnesHutSolver, "gpu-naive": NaiveSolver,
           "gpu-linear-bvh-4p": QuadrupoleLinearBVHSolver, "gpu-barnes-hut-4p": QuadrupoleBarnesHutSolver,
           "gpu-implicit-fmm-2p": ImplicitFMMSolver}


def select(solver_id: str):
    """MultiSolver::select(id) (src/NBody/Simulation/MultiSolver.h:79-87)."""
    try:
        return SOLVERS[solver_id]
    except KeyError:
        raise ValueError(f"unknown solver id {solver_id!r}") from None
