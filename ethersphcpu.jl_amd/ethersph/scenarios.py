"""Scene generators for the benchmark and test workloads.

Each follows a reference example's recipe (lattice generator, initial state,
wall layout, check predicate), with the sizes of the workload it names:

- ``collapse_dry_2d``  2D dry dam break (examples/CollapseDry/CollapseDry.jl:48-127):
  tank 4.0 x 3.0 m, water 1.0 x 2.0 m, dx = 0.02 -> 5,000 fluid + 1,518 wall.
- ``thermal_cavity_2d`` natural-convection cavity
  (examples/NaturalConvectionCavityNormalization/NaturalConvectionCavityNormalization.jl:17-135):
  unit cavity, dx = 0.005 -> 40,000 fluid, Ra = 1e5, Pr = 0.71, WendlandC2, h = 1.2 dx.
- ``dam_break_3d`` 3D dam break after examples/Cruchaga3D/Cruchaga3D.jl:69-160:
  box 3.2 x 1.0 x 1.6 m (z up), water 1.2 x 1.0 x 0.6 m, dx = 0.0089 ->
  1,013,040 fluid, 645,804 wall (five faces, no lid), WendlandC4, h = 1.3 dx.
- ``lid_driven_cavity_2d`` (examples/LidDrivenCavity/LidDrivenCavity.jl:10-156).

Lattice coordinates are computed with the reference's operation order,
x = (x0 + (j - 1) dx) + dx / 2, so that lattice pairs at exactly the cutoff
round the same way as in the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .models import CommonWeaklyCompressibleLiquidModel, ThermalWeaklyCompressibleLiquidModel
from .particles import LiquidParticles, WallParticles
from .solve import AABBCheck, Scene, make_desc, select_solver
from .sph_kernel import KERNELS, CubicSpline, SmoothKernel, WendlandC2, WendlandC4
from .time_discretization import DensityReinitializedFixedTimeIntervalForwardEuler


@dataclass
class Scenario:
    name: str
    dim: int
    kernel: object
    model: object
    td: DensityReinitializedFixedTimeIntervalForwardEuler
    fluid: LiquidParticles
    walls: list = field(default_factory=list)
    check: AABBCheck | None = None

    @property
    def solver(self) -> int:
        return select_solver(self.walls, self.model)

    @property
    def n_fluid(self) -> int:
        return len(self.fluid)

    @property
    def n_wall(self) -> int:
        return sum(len(w) for w in self.walls)

    def desc(self, precision: int = _lib.FP64, capacity: int = 0) -> _lib.SceneDesc:
        return make_desc(self.dim, self.kernel, self.model, self.td, self.solver, precision, self.check, capacity)

    def make_scene(self, precision: int = _lib.FP64, ctx=None, capacity: int = 0) -> Scene:
        return Scene(self.desc(precision, capacity), self.fluid, self.walls, ctx)


def rectangle(dx, x0, y0, width_number, height_number):
    """createRectangleParticles (CollapseDry.jl:51-58): row-major lattice."""
    j = np.arange(width_number, dtype=np.float64)
    i = np.arange(height_number, dtype=np.float64)
    xs = (x0 + j * dx) + dx / 2
    ys = (y0 + i * dx) + dx / 2
    X = np.empty((height_number * width_number, 2))
    X[:, 0] = np.tile(xs, height_number)
    X[:, 1] = np.repeat(ys, width_number)
    return X


def box(dx, x0, y0, z0, nx, ny, nz):
    """createBoxParticles (Cruchaga3D.jl:69-76): x fastest, then y, then z."""
    xs = (x0 + np.arange(nx, dtype=np.float64) * dx) + dx / 2
    ys = (y0 + np.arange(ny, dtype=np.float64) * dx) + dx / 2
    zs = (z0 + np.arange(nz, dtype=np.float64) * dx) + dx / 2
    X = np.empty((nx * ny * nz, 3))
    X[:, 0] = np.tile(xs, ny * nz)
    X[:, 1] = np.tile(np.repeat(ys, nx), nz)
    X[:, 2] = np.repeat(zs, nx * ny)
    return X


def _walls(kind, X, normal, gap, v=None, t=None, kappa=None):
    dim = X.shape[1]
    w = WallParticles(kind, dim, X.shape[0])
    w.x_vec_ = X
    w.normal_vec_ = np.tile(np.asarray(normal, dtype=np.float64), (X.shape[0], 1))
    w.gap_ = np.full(X.shape[0], gap)
    if v is not None:
        w.v_vec_ = np.tile(np.asarray(v, dtype=np.float64), (X.shape[0], 1))
    if t is not None:
        w.t_ = np.full(X.shape[0], t)
    if kappa is not None:
        w.kappa_ = np.full(X.shape[0], kappa)
    return w


def _ipow(x, n):
    y = np.ones_like(x)
    while n > 0:
        if n & 1:
            y = y * x
        x = x * x
        n >>= 1
    return y


def _update_pressure(f: LiquidParticles, m) -> None:
    """updatePressure! (UpdatePressure.jl:8-16) applied to the arrays."""
    rr = f.rho_ / m.rho_0_
    f.p_ = m.b_ * (_ipow(rr, m.gamma_) - 1.0) + m.p_0_
    f.c_ = m.c_0_ * np.sqrt(_ipow(rr, m.gamma_ - 1))


def collapse_dry_2d(dx=0.02, water_width=1.0, water_height=2.0, box_width=4.0, box_height=3.0,
                    gravity=9.81, rho_0=1000.0, gamma=1, mu_0=1e-3, h_factor=1.5, kernel=CubicSpline,
                    total_step=2000, reinit=30, check_interval=30):
    """2D dry dam break; defaults are the 5,000-particle workload (plain
    method, FP64, linear EOS, c0 = 10 sqrt(2 g H), dt = 0.1 h / c0)."""
    dim = 2
    R = kernel.radius_ratio * h_factor * dx
    sk = SmoothKernel(R, dim, kernel)
    c_0 = 10.0 * math.sqrt(2.0 * gravity * water_height)
    model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, 0.0, gamma, mu_0, [0.0, -gravity])
    dt = 0.1 * sk.h_ / c_0
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 100, check_interval, reinit)
    rows = round(water_height / dx)
    cols = round(water_width / dx)
    f = LiquidParticles(dim, rows * cols)
    f.x_vec_ = rectangle(dx, 0.0, 0.0, cols, rows)
    # CollapseDry.jl:65-72
    f.p_ = rho_0 * gravity * (water_height + dx - f.x_vec_[:, 1])
    f.rho_ = (f.p_ / model.b_ + 1.0) ** (1.0 / gamma) * rho_0
    _update_pressure(f, model)
    f.mass_ = f.rho_ * (dx * dx)
    f.gap_ = np.full(len(f), dx)
    t = 3
    bw = round(box_width / dx)
    bh = round(box_height / dx)
    groups = [
        _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0, 0.0 - t * dx, bw, t), [0.0, 1.0], dx),
        _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0 - t * dx, 0.0, t, bh), [1.0, 0.0], dx),
        _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0 + box_width, 0.0, t, bh), [-1.0, 0.0], dx),
        _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0 - t * dx, 0.0 - t * dx, t, t),
               np.array([1.0, 1.0]) / math.sqrt(2.0), dx),
        _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0 + box_width, 0.0 - t * dx, t, t),
               np.array([-1.0, 1.0]) / math.sqrt(2.0), dx),
    ]
    walls = WallParticles.concat(groups)
    check = AABBCheck(hi=[0.0 + box_width, 0.0 + box_height * 1.0])
    return Scenario("collapse_dry_2d", dim, sk, model, td, f, [walls], check)


def thermal_cavity_2d(dx=0.005, rayleigh=1e5, prandtl=0.71, c_0=4.0, h_factor=1.2, kernel=WendlandC2,
                      noise=1e-3, seed=0, total_step=20000, reinit=30):
    """Natural-convection cavity: hot left wall (T=1), cold right (T=0),
    adiabatic top/bottom, Boussinesq buoyancy, T0 = 0.5 + U(-noise, noise)."""
    dim = 2
    L = 1.0
    rho_0 = 1.0
    nu_0 = prandtl / math.sqrt(rayleigh)
    mu_0 = rho_0 * nu_0
    p_0 = 0.02 * rho_0 * c_0 * c_0
    gamma = 7
    alpha = nu_0 / prandtl
    g = 1.0
    t_left, t_right = 1.0, 0.0
    delta_t = t_left - t_right
    t_0 = (t_left + t_right) / 2
    kappa = 1.0
    cp = prandtl * kappa / mu_0
    beta = rayleigh * nu_0 * alpha / g / delta_t / (L * L * L)
    model = ThermalWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, [0.0, -g], t_0, kappa, cp, beta)
    R = kernel.radius_ratio * h_factor * dx
    sk = SmoothKernel(R, dim, kernel)
    dt = 0.1 * sk.h_ / c_0
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 100, 0, reinit)
    n = round(L / dx)
    f = LiquidParticles(dim, n * n, thermal=True)
    f.x_vec_ = rectangle(dx, 0.0, 0.0, n, n)
    rng = np.random.default_rng(seed)
    f.rho_ = np.full(len(f), rho_0)
    f.p_ = np.full(len(f), p_0)
    f.c_ = np.full(len(f), c_0)
    f.t_ = t_0 + rng.uniform(-noise, noise, len(f))
    f.mass_ = f.rho_ * (dx * dx)
    f.gap_ = np.full(len(f), dx)
    f.kappa_ = np.full(len(f), kappa)
    f.cp_ = np.full(len(f), cp)
    t = 3
    left = _walls(_lib.WALL_THERMOSTATIC, rectangle(dx, 0.0 - t * dx, 0.0 - t * dx, t, n + 2 * t), [1.0, 0.0], dx,
                  t=t_left, kappa=kappa)
    right = _walls(_lib.WALL_THERMOSTATIC, rectangle(dx, 0.0 + L, 0.0 - t * dx, t, n + 2 * t), [-1.0, 0.0], dx,
                   t=t_right, kappa=kappa)
    bottom = _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0, 0.0 - t * dx, n, t), [0.0, 1.0], dx)
    top = _walls(_lib.WALL_COMPULSIVE, rectangle(dx, 0.0, 0.0 + L, n, t), [0.0, -1.0], dx)
    walls = [WallParticles.concat([bottom, top]), WallParticles.concat([left, right])]
    return Scenario("thermal_cavity_2d", dim, sk, model, td, f, walls, None)


def dam_break_3d(dx=0.0089, box_size=(3.2, 1.0, 1.6), water_size=(1.2, 1.0, 0.6), gravity=9.81, rho_0=1000.0,
                 gamma=1, mu_0=1e-3, h_factor=1.3, kernel=WendlandC4, total_step=1000, reinit=30, lid=False):
    """3D dam break, z up; five wall faces (no lid unless lid=True), three layers."""
    dim = 3
    R = kernel.radius_ratio * h_factor * dx
    sk = SmoothKernel(R, dim, kernel)
    H = water_size[2]
    c_0 = 10.0 * math.sqrt(2.0 * gravity * H)
    model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, 0.0, gamma, mu_0, [0.0, 0.0, -gravity])
    dt = 0.1 * sk.h_ / c_0
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 100, 0, reinit)
    wa, wb, wc = (round(s / dx) for s in water_size)
    na, nb, nc = (round(s / dx) for s in box_size)
    f = LiquidParticles(dim, wa * wb * wc)
    f.x_vec_ = box(dx, 0.0, 0.0, 0.0, wa, wb, wc)
    # Cruchaga3D.jl:79-86 with the vertical axis z
    f.p_ = rho_0 * gravity * (H - f.x_vec_[:, 2])
    f.rho_ = (f.p_ / model.b_ + 1.0) ** (1.0 / gamma) * rho_0
    _update_pressure(f, model)
    f.mass_ = f.rho_ * (dx * dx * dx)
    f.gap_ = np.full(len(f), dx)
    t = 3
    C = _lib.WALL_COMPULSIVE
    groups = [
        _walls(C, box(dx, 0.0 - t * dx, 0.0 - t * dx, 0.0 - t * dx, na + 2 * t, nb + 2 * t, t), [0, 0, 1.0], dx),
        _walls(C, box(dx, 0.0 - t * dx, 0.0 - t * dx, 0.0, na + 2 * t, t, nc), [0, 1.0, 0], dx),
        _walls(C, box(dx, 0.0 - t * dx, 0.0 + nb * dx, 0.0, na + 2 * t, t, nc), [0, -1.0, 0], dx),
        _walls(C, box(dx, 0.0 - t * dx, 0.0, 0.0, t, nb, nc), [1.0, 0, 0], dx),
        _walls(C, box(dx, 0.0 + na * dx, 0.0, 0.0, t, nb, nc), [-1.0, 0, 0], dx),
    ]
    if lid:
        groups.append(_walls(C, box(dx, 0.0 - t * dx, 0.0 - t * dx, 0.0 + nc * dx, na + 2 * t, nb + 2 * t, t),
                             [0, 0, -1.0], dx))
    walls = WallParticles.concat(groups)
    return Scenario("dam_break_3d", dim, sk, model, td, f, [walls], None)


def lid_driven_cavity_2d(dx=0.01, reynolds=100.0, lid_velocity=1.0, kernel=WendlandC2, h_factor=1.5,
                         total_step=1000, reinit=30):
    """Lid-driven cavity (LidDrivenCavity.jl:10-156): StaticVelocityWall
    particles, lid at y = 1 moving at lid_velocity."""
    dim = 2
    L = 1.0
    rho_0 = 1.0
    mu_0 = L * rho_0 * lid_velocity / reynolds
    c_0 = 10.0 * lid_velocity
    p_0 = 0.001 * rho_0 * c_0 * c_0
    gamma = 7
    R = kernel.radius_ratio * h_factor * dx
    sk = SmoothKernel(R, dim, kernel if not isinstance(kernel, str) else KERNELS[kernel])
    model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, [0.0, 0.0])
    dt = 0.1 * sk.influence_radius_ / c_0
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 100, 0, reinit)
    n = round(L / dx)
    f = LiquidParticles(dim, n * n)
    f.x_vec_ = rectangle(dx, 0.0, 0.0, n, n)
    f.rho_ = np.full(len(f), rho_0)
    _update_pressure(f, model)
    f.c_ = np.full(len(f), c_0)
    f.mass_ = f.rho_ * (dx * dx)
    f.gap_ = np.full(len(f), dx)
    t = 3
    V = _lib.WALL_STATIC_VELOCITY
    s2 = math.sqrt(2.0)
    groups = [
        _walls(V, rectangle(dx, 0.0, 0.0 + L, n, t), [0.0, -1.0], dx, v=[lid_velocity, 0.0]),
        _walls(V, rectangle(dx, 0.0 - t * dx, 0.0, t, n), [1.0, 0.0], dx, v=[0.0, 0.0]),
        _walls(V, rectangle(dx, 0.0 + L, 0.0, t, n), [-1.0, 0.0], dx, v=[0.0, 0.0]),
        _walls(V, rectangle(dx, 0.0, 0.0 - t * dx, n, t), [0.0, 1.0], dx, v=[0.0, 0.0]),
        _walls(V, rectangle(dx, 0.0 - t * dx, 0.0 - t * dx, t, t), np.array([1.0, 1.0]) / s2, dx, v=[0.0, 0.0]),
        _walls(V, rectangle(dx, 0.0 + L, 0.0 - t * dx, t, t), np.array([-1.0, 1.0]) / s2, dx, v=[0.0, 0.0]),
        _walls(V, rectangle(dx, 0.0 - t * dx, 0.0 + L, t, t), np.array([1.0, -1.0]) / s2, dx, v=[lid_velocity, 0.0]),
        _walls(V, rectangle(dx, 0.0 + L, 0.0 + L, t, t), np.array([-1.0, -1.0]) / s2, dx, v=[lid_velocity, 0.0]),
    ]
    walls = WallParticles.concat(groups)
    return Scenario("lid_driven_cavity_2d", dim, sk, model, td, f, [walls], None)


def perturb(sc: Scenario, amplitude: float, seed: int = 0) -> Scenario:
    """Adds seeded uniform noise (fraction of the lattice gap) to the fluid
    positions and velocities: off-lattice states for neighbour-set checks."""
    rng = np.random.default_rng(seed)
    f = sc.fluid
    dx = float(f.gap_[0]) if len(f) else 1.0
    f.x_vec_ = f.x_vec_ + rng.uniform(-amplitude, amplitude, f.x_vec_.shape) * dx
    f.v_vec_ = f.v_vec_ + rng.uniform(-0.1, 0.1, f.v_vec_.shape)
    return sc
