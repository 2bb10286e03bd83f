"""Solve drivers: mirror of Solve/*.jl on top of the device library.

``solve(fluid, walls..., smooth_kernel, model, time_discretization, ...)``
picks the solver the way the reference's multiple dispatch does
(Solve/FluidWallSolve.jl:85-140, FluidVelocitySolve.jl:104-168,
FluidThermostaticWallSolve.jl:127-173, FluidCompulsiveThermostaticWallSolve.jl:154-233),
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
    "FluidInletOutletWall": _lib.SOLVER_FLUID_INLET_OUTLET_WALL,
}
PRECISIONS = {"FP64": _lib.FP64, "FP32_BULK": _lib.FP32_BULK}


class AABBCheck:
    """The ``check`` predicate as data: a fluid particle is kept iff, on every
    axis, not (x > hi) and not (x < lo) -- the form of the reference examples'
    closures (e.g. CollapseDry.jl:119-127) -- and, for every plane
    (normal, d) in ``planes`` (at most 8), not (normal . x > d).  Missing
    bounds are unbounded; together box and planes describe any convex
    keep-region bounded by planes."""

    def __init__(self, lo=None, hi=None, planes=()):
        self.lo = lo
        self.hi = hi
        self.planes = [(tuple(float(c) for c in n), float(d)) for n, d in planes]
        if len(self.planes) > 8:
            raise ValueError("at most 8 check planes")

    def bounds(self, dim):
        lo = np.full(3, -np.inf)
        hi = np.full(3, np.inf)
        if self.lo is not None:
            lo[:dim] = self.lo
        if self.hi is not None:
            hi[:dim] = self.hi
        return lo, hi


class InletOutletPlanes:
    """The inlet/outlet closures of FluidInletOutletWallSolve.jl:275-279 as
    data, in the form the Poiseuille examples use (PoiseuilleFlow2D.jl:146-151):
    along ``axis``, an inlet particle past ``inlet_enter`` is copied into the
    fluid and moved back by ``inlet_shift``; a fluid particle past
    ``outlet_enter`` moves to the outlet; an outlet particle past
    ``outlet_exit`` is removed."""

    def __init__(self, axis: int, inlet_enter: float, inlet_shift: float, outlet_enter: float,
                 outlet_exit: float):
        self.axis = int(axis)
        self.inlet_enter = float(inlet_enter)
        self.inlet_shift = float(inlet_shift)
        self.outlet_enter = float(outlet_enter)
        self.outlet_exit = float(outlet_exit)


class BufferProfile:
    """The examples' user_defined_function as data: buffer velocities along
    the axis set after every step at t = step * dt.

    kind CHANNEL_2D: the start-up channel series of PoiseuilleFlow2D.jl:28-35
    (``terms`` is the last n of the sum); PIPE_3D: the steady pipe profile of
    PoiseuilleFlow3D.jl:32-34.  ``outlet_profile`` sets the outlet's axial
    velocity too (2D example); the outlet's cross components are zeroed."""

    def __init__(self, kind: int, f: float, nu: float, l: float, center=(0.0, 0.0, 0.0), terms: int = 10,
                 outlet_profile: bool = True):
        self.kind = int(kind)
        self.f = float(f)
        self.nu = float(nu)
        self.l = float(l)
        self.center = tuple(float(c) for c in center) + (0.0,) * (3 - len(center))
        self.terms = int(terms)
        self.outlet_profile = bool(outlet_profile)


def select_solver(walls: list, model, buffers: bool = False) -> int:
    """Dispatch of solve! on particle-vector and model types."""
    thermal = isinstance(model, ThermalWeaklyCompressibleLiquidModel)
    kinds = [w.kind for w in walls]
    if buffers:  # FluidInletOutletWallSolve.jl:266-340
        if len(walls) == 1 and not thermal:
            return _lib.SOLVER_FLUID_INLET_OUTLET_WALL
        raise TypeError("the inlet/outlet solver takes one wall group and a common liquid model")
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
              check: AABBCheck | None = None, capacity: int = 0, domain=None,
              planes: InletOutletPlanes | None = None, profile: BufferProfile | None = None) -> _lib.SceneDesc:
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
        d.check_nplanes = len(check.planes)
        for q, (n, off) in enumerate(check.planes):
            for k in range(3):
                d.check_planes[q][k] = n[k] if k < len(n) else 0.0
            d.check_planes[q][3] = off
    d.fluid_capacity = capacity
    if domain is not None:
        d.domain_set = 1
        lo = np.zeros(3)
        hi = np.zeros(3)
        lo[:dim] = domain[0]
        hi[:dim] = domain[1]
        d.domain_min = _lib.D3(*lo)
        d.domain_max = _lib.D3(*hi)
    if planes is not None:
        d.io_axis = planes.axis
        d.inlet_enter = planes.inlet_enter
        d.inlet_shift = planes.inlet_shift
        d.outlet_enter = planes.outlet_enter
        d.outlet_exit = planes.outlet_exit
    if profile is not None:
        d.profile = profile.kind
        d.profile_terms = profile.terms
        d.outlet_profile = int(profile.outlet_profile)
        d.profile_f = profile.f
        d.profile_nu = profile.nu
        d.profile_l = profile.l
        d.profile_center = _lib.D3(*profile.center)
    return d


class Scene:
    """A scene resident on the device (esph_scene)."""

    _close_rank = 2

    def __init__(self, desc: _lib.SceneDesc, fluid: LiquidParticles, walls: list[WallParticles],
                 ctx: _lib.Context | None = None, buffers: tuple | None = None):
        """buffers: (inlet, outlet) LiquidParticles of the inlet/outlet solver."""
        self.ctx = ctx or _lib.default_context()
        self.L = self.ctx._L
        self.dim = desc.dim
        self.desc = desc
        self.thermal = desc.solver in (_lib.SOLVER_FLUID_THERMOSTATIC_WALL,
                                       _lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL)
        h = c_void_p()
        self.ctx.check(self.L.esph_scene_create(self.ctx.handle, byref(desc), byref(h)))
        self.handle = h
        self.ctx.adopt(self)
        try:
            self._set_walls(walls)
            if buffers is not None:
                self.set_buffers(*buffers)
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
        a, keep = fluid_arrays(f, self.thermal)
        self.ctx.check(self.L.esph_scene_set_fluid(self.handle, byref(a)))

    def set_buffers(self, inlet: LiquidParticles, outlet: LiquidParticles) -> None:
        """Inlet and outlet buffer particles (FluidInletOutletWallSolve.jl:8-20)."""
        a, k1 = fluid_arrays(inlet, False)
        b, k2 = fluid_arrays(outlet, False)
        self.ctx.check(self.L.esph_scene_set_buffers(self.handle, byref(a), byref(b)))

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

    def group_count(self, group: int) -> int:
        v = c_int64()
        self.ctx.check(self.L.esph_scene_group_count(self.handle, group, byref(v)))
        return int(v.value)

    def get(self, field: int, group: int = _lib.GROUP_FLUID) -> np.ndarray:
        n = self.group_count(group)
        w = self.dim if field in (_lib.FIELD_X, _lib.FIELD_V) else 1
        out = np.empty(max(n * w, 1))
        wr = c_int64()
        if group == _lib.GROUP_FLUID:
            self.ctx.check(self.L.esph_scene_get_fluid(self.handle, field, _lib.dptr(out), out.size, byref(wr)))
        else:
            self.ctx.check(self.L.esph_scene_get_group(self.handle, group, field, _lib.dptr(out), out.size,
                                                       byref(wr)))
        out = out[: wr.value]
        return out.reshape(n, w) if w > 1 else out

    def state(self, group: int = _lib.GROUP_FLUID) -> LiquidParticles:
        n = self.group_count(group)
        f = LiquidParticles(self.dim, n, self.thermal)
        f.x_vec_ = self.get(_lib.FIELD_X, group).reshape(n, self.dim)
        f.v_vec_ = self.get(_lib.FIELD_V, group).reshape(n, self.dim)
        for name, fid in (("rho_", _lib.FIELD_RHO), ("p_", _lib.FIELD_P), ("c_", _lib.FIELD_C),
                          ("mass_", _lib.FIELD_MASS), ("gap_", _lib.FIELD_GAP)):
            setattr(f, name, self.get(fid, group))
        if self.thermal:
            for name, fid in (("t_", _lib.FIELD_T), ("kappa_", _lib.FIELD_KAPPA), ("cp_", _lib.FIELD_CP)):
                setattr(f, name, self.get(fid, group))
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

    def walk_pairs(self, pairset: int = _lib.PAIRS_FLUID_FLUID) -> np.ndarray:
        """Sorted (m, 2) int32 directed pairs that the next step's pair
        passes decide to be neighbours (esph_scene_walk_pairs)."""
        m = c_int64()
        self.ctx.check(self.L.esph_scene_walk_pairs(self.handle, pairset, None, 0, byref(m)))
        out = np.empty((max(m.value, 1), 2), dtype=np.int32)
        self.ctx.check(self.L.esph_scene_walk_pairs(self.handle, pairset,
                                                    out.ctypes.data_as(ctypes.POINTER(c_int32)),
                                                    out.shape[0], byref(m)))
        return out[: m.value]

    def probe(self, points, field: int, none_value: float = float("nan")) -> np.ndarray:
        """SPH interpolation of a fluid field at points (n, dim):
        kernelValueInterpolation (DataIO/VTPToDataFrame.jl:77-142) on the
        device; (n, dim) for FIELD_V, (n,) otherwise."""
        P = _lib.f64(np.asarray(points, dtype=np.float64).reshape(-1, self.dim))
        w = self.dim if field == _lib.FIELD_V else 1
        out = np.empty(max(P.shape[0] * w, 1))
        self.ctx.check(self.L.esph_scene_probe(self.handle, field, P.shape[0], _lib.dptr(P), none_value,
                                               _lib.dptr(out)))
        out = out[: P.shape[0] * w]
        return out.reshape(-1, w) if w > 1 else out

    def probe_noalloc(self, points, field: int, none_value: float = float("nan")) -> np.ndarray:
        """probe() through esph_scene_probe_noalloc: no device allocation, at
        most 65,536 points per call."""
        P = _lib.f64(np.asarray(points, dtype=np.float64).reshape(-1, self.dim))
        w = self.dim if field == _lib.FIELD_V else 1
        out = np.empty(max(P.shape[0] * w, 1))
        self.ctx.check(self.L.esph_scene_probe_noalloc(self.handle, field, P.shape[0], _lib.dptr(P), none_value,
                                                       _lib.dptr(out)))
        out = out[: P.shape[0] * w]
        return out.reshape(-1, w) if w > 1 else out

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
        v = (c_int64 * 8)()
        self.ctx.check(self.L.esph_scene_stats(self.handle, v))
        return {"list_overflow": v[0], "list_width_ff": v[1], "list_width_fw": v[2], "cells": v[3],
                "outside_grid": v[4], "error_bits": v[5], "sorts": v[6], "reused_lists": v[7] > 0}

    def brick_stats(self) -> dict:
        """Brick walks of the scene: whether it walks bricks, their shape
        in cells, the bricks holding particles and the walks whose region
        exceeded the LDS staging."""
        v = (c_int64 * 6)()
        self.ctx.check(self.L.esph_scene_brick_stats(self.handle, v))
        return {"bricks": bool(v[0]), "brick_cells": [v[1], v[2], v[3]], "occupied": v[4], "spilled": v[5]}

    def enable_graph(self, on: bool = True) -> None:
        self.ctx.check(self.L.esph_scene_enable_graph(self.handle, int(on)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            # refused while a live batch holds the scene
            self.ctx.check(self.L.esph_scene_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fluid_arrays(f: LiquidParticles, thermal: bool):
    """esph_fluid_arrays borrowing f's columns; keep the returned arrays alive
    for the duration of the call."""
    a = _lib.FluidArrays()
    keep = {k: _lib.f64(getattr(f, k)) for k in ("x_vec_", "v_vec_", "rho_", "p_", "c_", "mass_", "gap_")}
    if thermal:
        for k in ("t_", "kappa_", "cp_"):
            keep[k] = _lib.f64(getattr(f, k))
    a.n = len(f)
    a.x, a.v = _lib.dptr(keep["x_vec_"]), _lib.dptr(keep["v_vec_"])
    a.rho, a.p, a.c = _lib.dptr(keep["rho_"]), _lib.dptr(keep["p_"]), _lib.dptr(keep["c_"])
    a.mass, a.gap = _lib.dptr(keep["mass_"]), _lib.dptr(keep["gap_"])
    if thermal:
        a.t, a.kappa, a.cp = _lib.dptr(keep["t_"]), _lib.dptr(keep["kappa_"]), _lib.dptr(keep["cp_"])
    return a, keep


def _as_fluid(fluid):
    if isinstance(fluid, LiquidParticles):
        return fluid
    return LiquidParticles.from_particles(fluid)


def _as_walls(w):
    if isinstance(w, WallParticles):
        return w
    return WallParticles.from_particles(w)


def _is_fluid_group(g) -> bool:
    if isinstance(g, LiquidParticles):
        return True
    return isinstance(g, (list, tuple)) and len(g) > 0 and not hasattr(g[0], "normal_vec_")


def solve(*args, check=None, precision: str = "FP64", ctx=None, on_output=None, inlet_outlet=None,
          user_defined_function=None):
    """solve!(fluid_particles, wall_particles..., smooth_kernel, model,
    time_discretization[, vtp_io]; check), and the inlet/outlet form
    solve!(fluid, inlet, outlet, walls, kernel, model, td[, vtp_io];
    inlet_outlet=InletOutletPlanes, user_defined_function=BufferProfile)
    (FluidInletOutletWallSolve.jl:266-340).

    Particle groups come first; the first SmoothKernel starts the trailing
    (kernel, model, time_discretization[, vtp_io]) arguments.  ``check`` must
    be an ``AABBCheck`` (arbitrary closures cannot run on the device).
    ``on_output(step, time, state)`` is called at output steps with the state
    read back.  With a VTPIO after the time discretisation, VTP files are
    written as the reference does, by a host thread that overlaps the device
    steps (vtp.AsyncVTPWriter).
    """
    k = next(i for i, a in enumerate(args) if isinstance(a, SmoothKernel_))
    groups = list(args[:k])
    kernel, model, td = args[k], args[k + 1], args[k + 2]
    vtp_io = args[k + 3] if len(args) > k + 3 else None
    if check is not None and not isinstance(check, AABBCheck):
        raise TypeError("check must be an AABBCheck: device steps cannot call host closures")
    io = len(groups) == 4 and all(_is_fluid_group(g) for g in groups[:3])
    if io and inlet_outlet is None:
        raise TypeError("the inlet/outlet solver needs inlet_outlet=InletOutletPlanes(...)")
    fluid_in = groups[0]
    fluid = _as_fluid(fluid_in)
    bufs_in = groups[1:3] if io else []
    bufs = [_as_fluid(b) for b in bufs_in]
    walls = [_as_walls(w) for w in groups[3 if io else 1:]]
    solver = select_solver(walls, model, buffers=io)
    desc = make_desc(fluid.dim, kernel, model, td, solver, PRECISIONS[precision], check,
                     planes=inlet_outlet if io else None, profile=user_defined_function if io else None)
    scene = Scene(desc, fluid, walls, ctx, buffers=tuple(bufs) if io else None)
    writer = None
    if vtp_io is not None:
        from .vtp import AsyncVTPWriter

        # writeVTP(0, 0.0, ...) then one file per output step, named by
        # div(step, output_interval) (FluidWallSolve.jl:94-95,135-137)
        writer = AsyncVTPWriter(vtp_io)
        writer.submit(0, 0.0, [fluid] + bufs + walls)
    try:
        total = td.total_step_
        done = 0
        while done < total:
            n = total - done
            if (on_output is not None or writer is not None) and td.output_interval_ > 0:
                n = min(n, td.output_interval_ - done % td.output_interval_)
            scene.step(n)
            done += n
            if (on_output is not None or writer is not None) and isOutputStep(done, td):
                st = scene.state()
                if on_output is not None:
                    on_output(done, done * td.dt_, st)
                if writer is not None:
                    bst = [scene.state(g) for g in (_lib.GROUP_INLET, _lib.GROUP_OUTLET)] if io else []
                    writer.submit(done // td.output_interval_, done * td.dt_, [st] + bst + walls)
        scene.sync()
        out = scene.state()
        ids = scene.ids()
        bouts = [scene.state(g) for g in (_lib.GROUP_INLET, _lib.GROUP_OUTLET)] if io else []
    finally:
        scene.close()
        if writer is not None:
            writer.close()
    if io:
        # the groups change size: fluid gains inlet copies and loses leavers,
        # the outlet gains them; results replace the callers' contents
        for dst, st in zip([fluid_in] + list(bufs_in), [out] + bouts):
            _write_group(dst, st)
        return None
    if isinstance(fluid_in, LiquidParticles):
        for f in ("x_vec_", "v_vec_") + LiquidParticles.SCALARS + (LiquidParticles.THERMAL if fluid.thermal else ()):
            setattr(fluid_in, f, getattr(out, f))
    else:
        kept = [fluid_in[i] for i in ids]
        fluid_in[:] = kept
        out.write_back(fluid_in)
    return None


def _write_group(dst, st: LiquidParticles) -> None:
    if isinstance(dst, LiquidParticles):
        for f in ("x_vec_", "v_vec_") + LiquidParticles.SCALARS:
            setattr(dst, f, getattr(st, f))
        return
    proto = type(dst[0]) if len(dst) else None
    if proto is None:
        raise ValueError("cannot grow an empty particle vector of unknown type")
    while len(dst) < len(st):
        dst.append(proto(st.dim))
    del dst[len(st):]
    st.write_back(dst)


# Python names cannot end in "!"; these aliases keep the reference's spelling
# recognisable.
solve_ = solve


class Batch:
    """Several resident scenes stepped together (esph_batch): one launch per
    (dim, kernel, precision) class and phase for all of them."""

    _close_rank = 1

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
        self.ctx.adopt(self)
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
