"""Solve drivers: mirror of Solve/*.jl on top of the device library.

``solve(fluid, walls..., smooth_kernel, model, time_discretization, ...)``
picks the solver the way the reference's multiple dispatch does
(Solve/FluidWallSolve.jl:101, FluidVelocitySolve.jl:104,
FluidThermostaticWallSolve.jl:127, FluidCompulsiveThermostaticWallSolve.jl:154),
uploads the particles once, advances ``total_step_`` steps on the GPU and
writes the state back into the caller's particles (deleting, in order, the
particles the check removed, as ``deleteat!`` does).

``Scene`` is the resident form: the particle state stays in HBM between
calls to ``step``; only ``get``/``state`` read it back.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import _lib
from .models import ThermalWeaklyCompressibleLiquidModel
from .particles import LiquidParticles, WallParticles
from .sph_kernel import SmoothKernel_
from .time_discretization import isOutputStep

SOLVER_NAMES = {
    "FluidWall": _lib.SOLVER_FLUID_WALL,
    "FluidVelocity": _lib.SOLVER_FLUID_VELOCITY,
    "FluidThermostaticWall": _lib.SOLVER_FLUID_THERMOSTATIC_WALL,
    "FluidCompulsiveThermostaticWall": _lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL,
}
PRECISIONS = {"FP64": _lib.FP64, "FP32_BULK": _lib.FP32_BULK}


class AABBCheck:
    """The ``check`` predicate as data: a fluid particle is kept iff, on every
    axis, not (x > hi) and not (x < lo) -- the form of the reference examples'
    closures (e.g. CollapseDry.jl:119-127).  Missing bounds are unbounded."""

    def __init__(self, lo=None, hi=None):
        self.lo = lo
        self.hi = hi

    def bounds(self, dim):
        lo = np.full(3, -np.inf)
        hi = np.full(3, np.inf)
        if self.lo is not None:
            lo[:dim] = self.lo
        if self.hi is not None:
            hi[:dim] = self.hi
        return lo, hi


def select_solver(walls: list, model) -> int:
    """Dispatch of solve! on particle-vector and model types."""
    thermal = isinstance(model, ThermalWeaklyCompressibleLiquidModel)
    kinds = [w.kind for w in walls]
    if len(walls) == 2 and thermal and kinds[1] == _lib.WALL_THERMOSTATIC:
        return _lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL
    if len(walls) == 1:
        if kinds[0] == _lib.WALL_STATIC_VELOCITY:
            return _lib.SOLVER_FLUID_VELOCITY
        if kinds[0] == _lib.WALL_THERMOSTATIC and thermal:
            return _lib.SOLVER_FLUID_THERMOSTATIC_WALL
        return _lib.SOLVER_FLUID_WALL
    if len(walls) == 0:
        return _lib.SOLVER_FLUID_WALL
    raise TypeError("no solve! method for this combination of particle groups and model")


def make_desc(dim, kernel: SmoothKernel_, model, td, solver: int, precision: int = _lib.FP64,
              check: AABBCheck | None = None, capacity: int = 0, domain=None) -> _lib.SceneDesc:
    d = _lib.SceneDesc()
    d.dim = dim
    d.kernel = kernel.kind
    d.influence_radius = kernel.influence_radius_
    d.solver = solver
    d.precision = precision
    d.rho_0 = model.rho_0_
    d.c_0 = model.c_0_
    d.p_0 = model.p_0_
    d.gamma = model.gamma_
    d.mu_0 = model.mu_0_
    bf = np.zeros(3)
    bf[:dim] = model.body_force_vec_[:dim]
    d.body_force = _lib.D3(*bf)
    if isinstance(model, ThermalWeaklyCompressibleLiquidModel):
        d.t_0 = model.t_0_
        d.kappa_0 = model.kappa_0_
        d.cp = model.cp_
        d.beta = model.beta_
    d.dt = td.dt_
    d.density_reinitialized_interval = td.density_reinitialized_interval_
    d.check_interval = td.check_interval_
    if check is not None:
        lo, hi = check.bounds(dim)
        d.check_enabled = 1
        d.check_min = _lib.D3(*lo)
        d.check_max = _lib.D3(*hi)
    d.fluid_capacity = capacity
    if domain is not None:
        d.domain_set = 1
        lo = np.zeros(3)
        hi = np.zeros(3)
        lo[:dim] = domain[0]
        hi[:dim] = domain[1]
        d.domain_min = _lib.D3(*lo)
        d.domain_max = _lib.D3(*hi)
    return d


class Scene:
    """A scene resident on the device (esph_scene)."""

    def __init__(self, desc: _lib.SceneDesc, fluid: LiquidParticles, walls: list[WallParticles],
                 ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.L = self.ctx._L
        self.dim = desc.dim
        self.desc = desc
        self.thermal = desc.solver in (_lib.SOLVER_FLUID_THERMOSTATIC_WALL,
                                       _lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL)
        h = c_void_p()
        self.ctx.check(self.L.esph_scene_create(self.ctx.handle, byref(desc), byref(h)))
        self.handle = h
        try:
            self._set_walls(walls)
            self.set_fluid(fluid)
        except Exception:
            self.close()
            raise

    def _set_walls(self, walls):
        keep = []
        arr = (_lib.WallArrays * max(len(walls), 1))()
        for q, w in enumerate(walls):
            x = _lib.f64(w.x_vec_)
            nrm = _lib.f64(w.normal_vec_)
            gap = _lib.f64(w.gap_)
            v = _lib.f64(w.v_vec_)
            t = _lib.f64(w.t_)
            k = _lib.f64(w.kappa_)
            keep += [x, nrm, gap, v, t, k]
            arr[q].kind = w.kind
            arr[q].n = len(w)
            arr[q].x, arr[q].normal, arr[q].gap = _lib.dptr(x), _lib.dptr(nrm), _lib.dptr(gap)
            arr[q].v, arr[q].t, arr[q].kappa = _lib.dptr(v), _lib.dptr(t), _lib.dptr(k)
        self.n_wall = sum(len(w) for w in walls)
        self.ctx.check(self.L.esph_scene_set_walls(self.handle, len(walls), arr))

    def set_fluid(self, f: LiquidParticles) -> None:
        a = _lib.FluidArrays()
        keep = {k: _lib.f64(getattr(f, k)) for k in ("x_vec_", "v_vec_", "rho_", "p_", "c_", "mass_", "gap_")}
        if self.thermal:
            for k in ("t_", "kappa_", "cp_"):
                keep[k] = _lib.f64(getattr(f, k))
        a.n = len(f)
        a.x, a.v = _lib.dptr(keep["x_vec_"]), _lib.dptr(keep["v_vec_"])
        a.rho, a.p, a.c = _lib.dptr(keep["rho_"]), _lib.dptr(keep["p_"]), _lib.dptr(keep["c_"])
        a.mass, a.gap = _lib.dptr(keep["mass_"]), _lib.dptr(keep["gap_"])
        if self.thermal:
            a.t, a.kappa, a.cp = _lib.dptr(keep["t_"]), _lib.dptr(keep["kappa_"]), _lib.dptr(keep["cp_"])
        self.ctx.check(self.L.esph_scene_set_fluid(self.handle, byref(a)))

    def step(self, nsteps: int = 1) -> None:
        self.ctx.check(self.L.esph_scene_step(self.handle, nsteps))

    def sync(self) -> None:
        self.ctx.check(self.L.esph_scene_sync(self.handle))

    @property
    def step_count(self) -> int:
        v = c_int64()
        self.ctx.check(self.L.esph_scene_step_count(self.handle, byref(v)))
        return int(v.value)

    def fluid_count(self) -> int:
        v = c_int64()
        self.ctx.check(self.L.esph_scene_fluid_count(self.handle, byref(v)))
        return int(v.value)

    def get(self, field: int) -> np.ndarray:
        n = self.fluid_count()
        w = self.dim if field in (_lib.FIELD_X, _lib.FIELD_V) else 1
        out = np.empty(max(n * w, 1))
        wr = c_int64()
        self.ctx.check(self.L.esph_scene_get_fluid(self.handle, field, _lib.dptr(out), out.size, byref(wr)))
        out = out[: wr.value]
        return out.reshape(n, w) if w > 1 else out

    def state(self) -> LiquidParticles:
        n = self.fluid_count()
        f = LiquidParticles(self.dim, n, self.thermal)
        f.x_vec_ = self.get(_lib.FIELD_X).reshape(n, self.dim)
        f.v_vec_ = self.get(_lib.FIELD_V).reshape(n, self.dim)
        for name, fid in (("rho_", _lib.FIELD_RHO), ("p_", _lib.FIELD_P), ("c_", _lib.FIELD_C),
                          ("mass_", _lib.FIELD_MASS), ("gap_", _lib.FIELD_GAP)):
            setattr(f, name, self.get(fid))
        if self.thermal:
            for name, fid in (("t_", _lib.FIELD_T), ("kappa_", _lib.FIELD_KAPPA), ("cp_", _lib.FIELD_CP)):
                setattr(f, name, self.get(fid))
        return f

    def ids(self) -> np.ndarray:
        return self.get(_lib.FIELD_ID).astype(np.int64)

    def neighbours(self, pairset: int = _lib.PAIRS_FLUID_FLUID) -> np.ndarray:
        """Sorted (m, 2) int32 pair list the next step will use."""
        m = c_int64()
        self.ctx.check(self.L.esph_scene_neighbours(self.handle, pairset, None, 0, byref(m)))
        out = np.empty((max(m.value, 1), 2), dtype=np.int32)
        self.ctx.check(self.L.esph_scene_neighbours(self.handle, pairset,
                                                    out.ctypes.data_as(ctypes.POINTER(c_int32)),
                                                    out.shape[0], byref(m)))
        return out[: m.value]

    def enable_timing(self, on: bool = True) -> None:
        self.ctx.check(self.L.esph_scene_enable_timing(self.handle, int(on)))

    def phase_times(self):
        """Average device ms per step: sort, continuity, forces, Shepard (per
        Shepard step), reserved; and the number of steps timed."""
        ms = (c_double * 5)()
        n = c_int64()
        self.ctx.check(self.L.esph_scene_phase_times(self.handle, ms, byref(n)))
        return {"sort": ms[0], "continuity": ms[1], "forces": ms[2], "shepard": ms[3], "steps": n.value}

    def stats(self) -> dict:
        v = (c_int64 * 4)()
        self.ctx.check(self.L.esph_scene_stats(self.handle, v))
        return {"list_overflow": v[0], "list_width_ff": v[1], "list_width_fw": v[2], "cells": v[3]}

    def enable_graph(self, on: bool = True) -> None:
        self.ctx.check(self.L.esph_scene_enable_graph(self.handle, int(on)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.L.esph_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_fluid(fluid):
    if isinstance(fluid, LiquidParticles):
        return fluid
    return LiquidParticles.from_particles(fluid)


def _as_walls(w):
    if isinstance(w, WallParticles):
        return w
    return WallParticles.from_particles(w)


def solve(*args, check=None, precision: str = "FP64", ctx=None, on_output=None):
    """solve!(fluid_particles, wall_particles..., smooth_kernel, model,
    time_discretization[, vtp_io]; check).

    Particle groups come first; the first SmoothKernel starts the trailing
    (kernel, model, time_discretization[, vtp_io]) arguments.  ``check`` must
    be an ``AABBCheck`` (arbitrary closures cannot run on the device).
    ``on_output(step, time, state)`` is called at output steps with the state
    read back (VTP writing itself is not part of this library).
    """
    k = next(i for i, a in enumerate(args) if isinstance(a, SmoothKernel_))
    groups = list(args[:k])
    kernel, model, td = args[k], args[k + 1], args[k + 2]
    if check is not None and not isinstance(check, AABBCheck):
        raise TypeError("check must be an AABBCheck: device steps cannot call host closures")
    fluid_in = groups[0]
    fluid = _as_fluid(fluid_in)
    walls = [_as_walls(w) for w in groups[1:]]
    solver = select_solver(walls, model)
    desc = make_desc(fluid.dim, kernel, model, td, solver, PRECISIONS[precision], check)
    scene = Scene(desc, fluid, walls, ctx)
    try:
        total = td.total_step_
        done = 0
        while done < total:
            n = total - done
            if on_output is not None and td.output_interval_ > 0:
                n = min(n, td.output_interval_ - done % td.output_interval_)
            scene.step(n)
            done += n
            if on_output is not None and isOutputStep(done, td):
                on_output(done, done * td.dt_, scene.state())
        scene.sync()
        out = scene.state()
        ids = scene.ids()
    finally:
        scene.close()
    if isinstance(fluid_in, LiquidParticles):
        for f in ("x_vec_", "v_vec_") + LiquidParticles.SCALARS + (LiquidParticles.THERMAL if fluid.thermal else ()):
            setattr(fluid_in, f, getattr(out, f))
    else:
        kept = [fluid_in[i] for i in ids]
        fluid_in[:] = kept
        out.write_back(fluid_in)
    return None


# Python names cannot end in "!"; these aliases keep the reference's spelling
# recognisable.
solve_ = solve


class Batch:
    """Several resident scenes stepped together (esph_batch): one launch per
    (dim, kernel, precision) class and phase for all of them."""

    def __init__(self, scenes: list, graph: bool = True):
        if not scenes:
            raise ValueError("empty batch")
        self.scenes = list(scenes)
        self.ctx = self.scenes[0].ctx
        self.L = self.ctx._L
        arr = (c_void_p * len(self.scenes))(*[s.handle.value for s in self.scenes])
        h = c_void_p()
        self.ctx.check(self.L.esph_batch_create(self.ctx.handle, len(self.scenes), arr, byref(h)))
        self.handle = h
        if graph:
            self.ctx.check(self.L.esph_batch_enable_graph(self.handle, 1))

    def step(self, nsteps: int = 1) -> None:
        self.ctx.check(self.L.esph_batch_step(self.handle, nsteps))

    def sync(self) -> None:
        self.ctx.check(self.L.esph_batch_sync(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.L.esph_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
