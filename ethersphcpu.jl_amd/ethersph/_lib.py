"""ctypes binding of the C ABI in include/ethersph.h.

The shared library ``libethersph_hip.so`` (built next to this package by
``make -C ethersphcpu.jl_amd``) is loaded lazily on first use; importing this
module only defines the ABI structures.  There is no fallback: every compute
call goes through the HIP library, and a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libethersph_hip.so")

# ---- enums (include/ethersph.h) ------------------------------------------------
ESPH_OK = 0
ERRORS = {1: "ESPH_ERR_ARG", 2: "ESPH_ERR_HIP", 3: "ESPH_ERR_BUDGET", 4: "ESPH_ERR_NAN",
          5: "ESPH_ERR_STATE", 6: "ESPH_ERR_CAPACITY", 7: "ESPH_ERR_NO_DEVICE"}

CUBIC_SPLINE, GAUSSIAN, WENDLAND_C2, WENDLAND_C4 = 0, 1, 2, 3
SOLVER_FLUID_WALL, SOLVER_FLUID_VELOCITY, SOLVER_FLUID_THERMOSTATIC_WALL, \
    SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL, SOLVER_FLUID_INLET_OUTLET_WALL = 0, 1, 2, 3, 4
PROFILE_NONE, PROFILE_CHANNEL_2D, PROFILE_PIPE_3D = 0, 1, 2
GROUP_FLUID, GROUP_INLET, GROUP_OUTLET = 0, 1, 2
FP64, FP32_BULK = 0, 1
WALL_COMPULSIVE, WALL_STATIC_VELOCITY, WALL_THERMOSTATIC = 0, 1, 2
FIELD_X, FIELD_V, FIELD_RHO, FIELD_P, FIELD_C, FIELD_MASS, FIELD_GAP, FIELD_T, FIELD_KAPPA, \
    FIELD_CP, FIELD_ID = range(11)
PAIRS_FLUID_FLUID, PAIRS_FLUID_WALL = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

D3 = c_double * 3


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("dim", c_int32), ("kernel", c_int32), ("influence_radius", c_double),
        ("solver", c_int32), ("precision", c_int32),
        ("rho_0", c_double), ("c_0", c_double), ("p_0", c_double), ("gamma", c_int32),
        ("mu_0", c_double), ("body_force", D3),
        ("t_0", c_double), ("kappa_0", c_double), ("cp", c_double), ("beta", c_double),
        ("dt", c_double), ("density_reinitialized_interval", c_int64), ("check_interval", c_int64),
        ("check_enabled", c_int32), ("check_min", D3), ("check_max", D3),
        ("fluid_capacity", c_int64), ("domain_set", c_int32), ("domain_min", D3), ("domain_max", D3),
        ("io_axis", c_int32), ("inlet_enter", c_double), ("inlet_shift", c_double), ("outlet_enter", c_double),
        ("outlet_exit", c_double), ("profile", c_int32), ("profile_terms", c_int32), ("outlet_profile", c_int32),
        ("profile_f", c_double), ("profile_nu", c_double), ("profile_l", c_double), ("profile_center", D3),
        ("verlet_skin", c_double),
        ("check_nplanes", c_int32), ("check_planes", (c_double * 4) * 8),
    ]


DP = POINTER(c_double)


class FluidArrays(ctypes.Structure):
    _fields_ = [("n", c_int64), ("x", DP), ("v", DP), ("rho", DP), ("p", DP), ("c", DP),
                ("mass", DP), ("gap", DP), ("t", DP), ("kappa", DP), ("cp", DP)]


class WallArrays(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("n", c_int64), ("x", DP), ("normal", DP), ("gap", DP),
                ("v", DP), ("t", DP), ("kappa", DP)]


class EsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libethersph_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run make -C {PKG_ROOT})")
    L = ctypes.CDLL(LIB_PATH)
    vp = c_void_p
    sig = {
        "esph_abi_version": ([], c_int32),
        "esph_ctx_create": ([c_int32, c_uint64, POINTER(vp)], c_int32),
        "esph_ctx_destroy": ([vp], c_int32),
        "esph_last_error": ([vp], c_char_p),
        "esph_ctx_memory_used": ([vp, POINTER(c_uint64)], c_int32),
        "esph_ctx_tuned_pass": ([vp, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(c_int32), POINTER(c_int32),
                                 POINTER(c_float)], c_int32),
        "esph_ctx_set_tuning": ([vp, c_int32, c_int32], c_int32),
        "esph_ctx_set_workgroup": ([vp, c_int32], c_int32),
        "esph_ctx_set_brick_walk": ([vp, c_int32], c_int32),
        "esph_ctx_set_wall_lanes": ([vp, c_int32], c_int32),
        "esph_ctx_tuned_workgroup": ([vp, c_int32, c_int32, c_int32, c_int32, POINTER(c_int32), POINTER(c_float)],
                                     c_int32),
        "esph_ctx_tuned_depth": ([vp, c_int32, c_int32, c_int32, c_int32, POINTER(c_int32), POINTER(c_float)],
                                 c_int32),
        "esph_kernel_eval": ([vp, c_int32, c_int32, c_double, c_int64, DP, DP, DP], c_int32),
        "esph_neighbor_list": ([vp, c_int32, c_double, c_int64, DP, c_int64, DP, POINTER(c_int32), c_int64,
                                POINTER(c_int64)], c_int32),
        "esph_scene_create": ([vp, POINTER(SceneDesc), POINTER(vp)], c_int32),
        "esph_scene_destroy": ([vp], c_int32),
        "esph_scene_set_walls": ([vp, c_int32, POINTER(WallArrays)], c_int32),
        "esph_scene_set_fluid": ([vp, POINTER(FluidArrays)], c_int32),
        "esph_scene_set_buffers": ([vp, POINTER(FluidArrays), POINTER(FluidArrays)], c_int32),
        "esph_scene_group_count": ([vp, c_int32, POINTER(c_int64)], c_int32),
        "esph_scene_get_group": ([vp, c_int32, c_int32, DP, c_int64, POINTER(c_int64)], c_int32),
        "esph_scene_step": ([vp, c_int64], c_int32),
        "esph_scene_sync": ([vp], c_int32),
        "esph_scene_step_count": ([vp, POINTER(c_int64)], c_int32),
        "esph_scene_fluid_count": ([vp, POINTER(c_int64)], c_int32),
        "esph_scene_get_fluid": ([vp, c_int32, DP, c_int64, POINTER(c_int64)], c_int32),
        "esph_scene_neighbours": ([vp, c_int32, POINTER(c_int32), c_int64, POINTER(c_int64)], c_int32),
        "esph_scene_walk_pairs": ([vp, c_int32, POINTER(c_int32), c_int64, POINTER(c_int64)], c_int32),
        "esph_scene_probe": ([vp, c_int32, c_int64, DP, c_double, DP], c_int32),
        "esph_scene_probe_noalloc": ([vp, c_int32, c_int64, DP, c_double, DP], c_int32),
        "esph_scene_enable_timing": ([vp, c_int32], c_int32),
        "esph_scene_phase_times": ([vp, DP, POINTER(c_int64)], c_int32),
        "esph_scene_stats": ([vp, POINTER(c_int64)], c_int32),
        "esph_scene_brick_stats": ([vp, POINTER(c_int64)], c_int32),
        "esph_scene_enable_graph": ([vp, c_int32], c_int32),
        "esph_batch_create": ([vp, c_int32, POINTER(vp), POINTER(vp)], c_int32),
        "esph_batch_destroy": ([vp], c_int32),
        "esph_batch_enable_graph": ([vp, c_int32], c_int32),
        "esph_batch_step": ([vp, c_int64], c_int32),
        "esph_batch_sync": ([vp], c_int32),
        "esph_server_create": ([vp, c_int32, POINTER(vp)], c_int32),
        "esph_server_destroy": ([vp], c_int32),
        "esph_server_submit": ([vp, POINTER(SceneDesc), c_int32, POINTER(WallArrays), POINTER(FluidArrays),
                                POINTER(FluidArrays), POINTER(FluidArrays), c_int64, POINTER(c_int64)], c_int32),
        "esph_server_advance": ([vp, c_int64, POINTER(c_int64)], c_int32),
        "esph_server_status": ([vp, c_int64, POINTER(c_int32), POINTER(c_int64)], c_int32),
        "esph_server_result": ([vp, c_int64, c_int32, DP, c_int64, POINTER(c_int64)], c_int32),
        "esph_server_release": ([vp, c_int64], c_int32),
        "esph_server_stats": ([vp, DP], c_int32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def dptr(a: np.ndarray | None):
    """Pointer to a C-contiguous float64 array (or NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float64"
    return a.ctypes.data_as(DP)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context (esph_ctx): a HIP stream and an HBM budget."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._L = lib()
        h = c_void_p()
        rc = self._L.esph_ctx_create(device, hbm_budget_bytes, ctypes.byref(h))
        if rc != ESPH_OK:
            raise EsphError(rc, "esph_ctx_create failed (no GPU?)")
        self.handle = h
        # scenes, batches and servers of this context, closed before it
        self._children = weakref.WeakSet()

    def adopt(self, obj) -> None:
        self._children.add(obj)

    def check(self, rc: int) -> None:
        if rc != ESPH_OK:
            raise EsphError(rc, self._L.esph_last_error(self.handle).decode())

    def set_tuning(self, tune: bool = True, depth: int = 2) -> None:
        """Gather depth (and workgroup size) of the pair passes: tuned per
        scene class at its first step (default), or fixed (depth 1, 2 or 4;
        workgroup size from set_workgroup)."""
        self.check(self._L.esph_ctx_set_tuning(self.handle, 1 if tune else 0, depth))

    def set_workgroup(self, threads: int = 256) -> None:
        """Workgroup size of the pair passes (64, 128 or 256 threads) when
        tuning is off."""
        self.check(self._L.esph_ctx_set_workgroup(self.handle, threads))

    def set_brick_walk(self, enable: bool = True) -> None:
        """Brick walks (LDS-staged neighbour regions) for the 3D FP32_BULK
        scenes that allow them (default off); applies to scenes uploaded
        afterwards.  Results are identical either way."""
        self.check(self._L.esph_ctx_set_brick_walk(self.handle, 1 if enable else 0))

    def set_wall_lanes(self, enable: bool = True) -> None:
        """3D force-pass wall walks in each particle's own lane (True) or
        shared out over the wavefront's lanes (default; 2D scenes always walk
        per lane); applies to scenes uploaded afterwards.  The two differ in
        the rounding of the wall sums only."""
        self.check(self._L.esph_ctx_set_wall_lanes(self.handle, 1 if enable else 0))

    def tuned_workgroup(self, dim: int, kernel: int, precision: int, thermal_mix: int = 1):
        """(threads, [ms at 64, 128, 256 threads]) of a tuned class, or None."""
        t = c_int32()
        ms = (c_float * 3)()
        rc = self._L.esph_ctx_tuned_workgroup(self.handle, dim, kernel, precision, thermal_mix, ctypes.byref(t), ms)
        if rc != ESPH_OK:
            return None
        return int(t.value), [float(v) for v in ms]

    def tuned_depth(self, dim: int, kernel: int, precision: int, thermal_mix: int = 1):
        """(depth, [ms at depth 1, 2, 4]) of a tuned class, or None."""
        d = c_int32()
        ms = (c_float * 3)()
        rc = self._L.esph_ctx_tuned_depth(self.handle, dim, kernel, precision, thermal_mix, ctypes.byref(d), ms)
        if rc != ESPH_OK:
            return None
        return int(d.value), [float(v) for v in ms]

    def tuned_pass(self, dim: int, kernel: int, precision: int, pass_: int, thermal_mix: int = 1):
        """(depth, threads, ms9) chosen for a pass (0 forces, 1 continuity, 2
        Shepard) of a class, or None before it is tuned."""
        d, t, ms = c_int32(), c_int32(), (c_float * 9)()
        rc = self._L.esph_ctx_tuned_pass(self.handle, dim, kernel, precision, thermal_mix, pass_, ctypes.byref(d),
                                         ctypes.byref(t), ms)
        if rc == 5:
            return None
        self.check(rc)
        return d.value, t.value, list(ms)

    def memory_used(self) -> int:
        v = c_uint64()
        self.check(self._L.esph_ctx_memory_used(self.handle, ctypes.byref(v)))
        return int(v.value)

    def close(self) -> None:
        if self.handle:
            # servers and batches first: they reference scenes
            kids = list(getattr(self, "_children", ()))
            for rank in (0, 1, 2):
                for k in kids:
                    if getattr(k, "_close_rank", 2) == rank:
                        k.close()
            self._L.esph_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
