"""ctypes wrapper of the FP64 CPU oracle (libesph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_double, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libesph_oracle.so")

_L = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        # the ABI structures are shared with the product's binding
        from ethersph._lib import FluidArrays, SceneDesc, WallArrays

        L = ctypes.CDLL(LIB)
        DP = POINTER(c_double)
        vp = c_void_p
        sig = {
            "eo_kernel_eval": ([c_int32, c_int32, c_double, c_int64, DP, DP, DP, DP], c_int32),
            "eo_neighbor_list": ([c_int32, c_double, c_int64, DP, c_int64, DP, c_int32, c_int32, POINTER(c_int32),
                                  c_int64], c_int64),
            "eo_scene_new": ([POINTER(SceneDesc)], vp),
            "eo_scene_free": ([vp], None),
            "eo_scene_set_walls": ([vp, c_int32, POINTER(WallArrays)], c_int32),
            "eo_scene_set_fluid": ([vp, POINTER(FluidArrays)], c_int32),
            "eo_scene_set_buffers": ([vp, POINTER(FluidArrays), POINTER(FluidArrays)], c_int32),
            "eo_scene_group_count": ([vp, c_int32], c_int64),
            "eo_scene_get_group": ([vp, c_int32, c_int32, DP], c_int32),
            "eo_scene_step": ([vp, c_int64, c_int32], c_int32),
            "eo_scene_fluid_count": ([vp], c_int64),
            "eo_scene_step_count": ([vp], c_int64),
            "eo_scene_get_fluid": ([vp, c_int32, DP], c_int32),
            "eo_scene_neighbours": ([vp, c_int32, POINTER(c_int32), c_int64], c_int64),
            "eo_scene_probe": ([vp, c_int32, c_int64, DP, c_double, DP], c_int32),
        }
        for n, (a, r) in sig.items():
            getattr(L, n).argtypes = a
            getattr(L, n).restype = r
        _L = L
    return _L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(POINTER(c_double))


def kernel_eval(kind: int, dim: int, R: float, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    w = np.empty_like(r)
    dw = np.empty_like(r)
    p3 = np.empty(3)
    assert lib().eo_kernel_eval(kind, dim, R, r.size, _p(r), _p(w), _p(dw), _p(p3)) == 0
    return w, dw, {"h": p3[0], "sigma": p3[1], "kernel_0": p3[2]}


def neighbor_list(x, cutoff, y=None, brute=False, nthreads=1) -> np.ndarray:
    X = np.ascontiguousarray(x, dtype=np.float64)
    Y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    ny = 0 if Y is None else Y.shape[0]
    L = lib()
    m = L.eo_neighbor_list(X.shape[1], cutoff, X.shape[0], _p(X), ny, _p(Y), int(brute), nthreads, None, 0)
    assert m >= 0
    out = np.empty((max(m, 1), 2), dtype=np.int32)
    L.eo_neighbor_list(X.shape[1], cutoff, X.shape[0], _p(X), ny, _p(Y), int(brute), nthreads,
                       out.ctypes.data_as(POINTER(c_int32)), m)
    return out[:m]


class OracleScene:
    """FP64 CPU restatement of one scene (same descriptor and arrays as the
    device Scene)."""

    def __init__(self, desc, fluid, walls, buffers=None):
        from ethersph._lib import WallArrays

        self.L = lib()
        self.dim = desc.dim
        self.thermal = desc.solver in (2, 3)
        self.h = self.L.eo_scene_new(byref(desc))
        assert self.h, "bad scene descriptor"
        arr = (WallArrays * max(len(walls), 1))()
        self._keep = []
        for q, w in enumerate(walls):
            a = [np.ascontiguousarray(getattr(w, f), dtype=np.float64)
                 for f in ("x_vec_", "normal_vec_", "gap_", "v_vec_", "t_", "kappa_")]
            self._keep += a
            arr[q].kind = w.kind
            arr[q].n = len(w)
            arr[q].x, arr[q].normal, arr[q].gap, arr[q].v, arr[q].t, arr[q].kappa = (_p(v) for v in a)
        assert self.L.eo_scene_set_walls(self.h, len(walls), arr) == 0
        if buffers is not None:
            a, k1 = _fluid_arrays(buffers[0], False)
            b, k2 = _fluid_arrays(buffers[1], False)
            assert self.L.eo_scene_set_buffers(self.h, byref(a), byref(b)) == 0
        f, keep = _fluid_arrays(fluid, self.thermal)
        assert self.L.eo_scene_set_fluid(self.h, byref(f)) == 0

    def step(self, n: int, nthreads: int = 1) -> None:
        self.L.eo_scene_step(self.h, n, nthreads)

    def fluid_count(self) -> int:
        return int(self.L.eo_scene_fluid_count(self.h))

    def group_count(self, group: int) -> int:
        return int(self.L.eo_scene_group_count(self.h, group))

    def get(self, field: int, group: int = 0) -> np.ndarray:
        n = self.group_count(group)
        w = self.dim if field in (0, 1) else 1
        out = np.empty(max(n * w, 1))
        if group == 0:
            assert self.L.eo_scene_get_fluid(self.h, field, _p(out)) == 0
        else:
            assert self.L.eo_scene_get_group(self.h, group, field, _p(out)) == 0
        out = out[: n * w]
        return out.reshape(n, w) if w > 1 else out

    def probe(self, points, field: int, none_value: float = float("nan")) -> np.ndarray:
        P = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, self.dim))
        w = self.dim if field == 1 else 1
        out = np.empty(max(P.shape[0] * w, 1))
        assert self.L.eo_scene_probe(self.h, field, P.shape[0], _p(P), none_value, _p(out)) == 0
        out = out[: P.shape[0] * w]
        return out.reshape(-1, w) if w > 1 else out

    def neighbours(self, pairset: int = 0) -> np.ndarray:
        m = self.L.eo_scene_neighbours(self.h, pairset, None, 0)
        out = np.empty((max(m, 1), 2), dtype=np.int32)
        self.L.eo_scene_neighbours(self.h, pairset, out.ctypes.data_as(POINTER(c_int32)), m)
        return out[:m]

    def close(self) -> None:
        if self.h:
            self.L.eo_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fluid_arrays(fluid, thermal):
    from ethersph._lib import FluidArrays

    f = FluidArrays()
    keep = {k: np.ascontiguousarray(getattr(fluid, k), dtype=np.float64)
            for k in ("x_vec_", "v_vec_", "rho_", "p_", "c_", "mass_", "gap_")}
    if thermal:
        for k in ("t_", "kappa_", "cp_"):
            keep[k] = np.ascontiguousarray(getattr(fluid, k), dtype=np.float64)
    f.n = len(fluid)
    f.x, f.v, f.rho, f.p, f.c = (_p(keep[k]) for k in ("x_vec_", "v_vec_", "rho_", "p_", "c_"))
    f.mass, f.gap = _p(keep["mass_"]), _p(keep["gap_"])
    if thermal:
        f.t, f.kappa, f.cp = _p(keep["t_"]), _p(keep["kappa_"]), _p(keep["cp_"])
    return f, keep


def from_scenario(sc, precision: int = 0) -> OracleScene:
    return OracleScene(sc.desc(precision), sc.fluid, sc.walls, getattr(sc, "buffers", None))
