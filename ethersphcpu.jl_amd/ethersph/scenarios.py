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
- ``poiseuille_2d`` channel with inlet/outlet buffers
  (examples/PoiseuilleFlow2D/PoiseuilleFlow2D.jl:11-169): l = 1e-2, dr = l / 50
  -> 10,000 fluid, 150 + 150 buffer, 1,272 wall; WendlandC2, R = 3 dr.
- ``poiseuille_3d`` pipe with inlet/outlet buffers
  (examples/PoiseuilleFlow3D/PoiseuilleFlow3D.jl:11-158): cylinder-ring
  lattice, dr = l / 40.

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
from .solve import AABBCheck, BufferProfile, InletOutletPlanes, Scene, make_desc, select_solver
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
    buffers: tuple | None = None  # (inlet, outlet) of the inlet/outlet solver
    planes: InletOutletPlanes | None = None
    profile: BufferProfile | None = None

    @property
    def solver(self) -> int:
        return select_solver(self.walls, self.model, buffers=self.buffers is not None)

    @property
    def n_fluid(self) -> int:
        return len(self.fluid)

    @property
    def n_wall(self) -> int:
        return sum(len(w) for w in self.walls)

    @property
    def n_particles(self) -> int:
        """Fluid plus buffer particles (the device particle arrays)."""
        return self.n_fluid + (sum(len(b) for b in self.buffers) if self.buffers else 0)

    def desc(self, precision: int = _lib.FP64, capacity: int = 0) -> _lib.SceneDesc:
        return make_desc(self.dim, self.kernel, self.model, self.td, self.solver, precision, self.check, capacity,
                         planes=self.planes, profile=self.profile)

    def make_scene(self, precision: int = _lib.FP64, ctx=None, capacity: int = 0) -> Scene:
        return Scene(self.desc(precision, capacity), self.fluid, self.walls, ctx, buffers=self.buffers)


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


def viscous_dt(dt, h, nu):
    """The examples' acoustic step 0.1 R / c0 (or 0.1 h / c0), capped by the
    explicit viscous limit 0.125 h^2 / nu: scenes rescaled to finer spacing
    or lower Reynolds number than the examples' own would otherwise take
    steps the forward-Euler viscous term cannot sustain.  At the examples'
    own parameters the acoustic step is the smaller one and is kept."""
    return min(dt, 0.125 * h * h / nu) if nu > 0 else dt


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
    dt = viscous_dt(dt, sk.h_, model.mu_0_ / model.rho_0_)
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


def _channel_ux(y, t, f, nu, l, terms):
    """PoiseuilleFlow2D.jl:28-35 (y measured from the bottom wall)."""
    ux = f / 2.0 / nu * y * (l - y)
    for n in range(terms + 1):
        o = float(2 * n + 1)
        ux = ux - 4.0 * f * (l * l) / nu / (math.pi ** 3) / (o * o * o) * np.sin(math.pi * y / l * o) * np.exp(
            -(o * o) * (math.pi * math.pi) * nu * t / (l * l))
    return ux


def poiseuille_2d(rows=50, reynolds=None, l=1e-2, nu=1e-6, rho_0=1000.0, f=8e-5, kernel=WendlandC2,
                  radius_factor=3.0, length_factor=4, terms=10, total_step=1000, reinit=5, check=5):
    """2D Poiseuille channel with inlet and outlet buffers
    (PoiseuilleFlow2D.jl:11-169): ``rows`` lattice rows across the channel
    (dr = l / rows; the example has 50), ``length_factor * l`` long, three
    buffer columns at each end, three compulsive-wall layers above and below
    that reach three columns past the buffers.  The fluid starts on the
    analytic start-up profile at t = 0.  ``reynolds`` (Re = u_max l / nu),
    when given, replaces the body force: f = 8 nu^2 Re / l^3."""
    dim = 2
    dr = l / rows
    gap = dr
    if reynolds is not None:
        f = 8.0 * nu * nu * reynolds / (l * l * l)
    R = radius_factor * dr
    sk = SmoothKernel(R, dim, kernel if not isinstance(kernel, str) else KERNELS[kernel])
    ux_max = f / 2 / nu / 4 * l ** 2
    c_0 = 10 * ux_max
    p_0 = 0.02 * rho_0 * c_0 ** 2
    mu_0 = nu * rho_0
    gamma = 7
    model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, [f, 0.0])
    dt = 0.1 * sk.influence_radius_ / c_0
    dt = viscous_dt(dt, sk.h_, model.mu_0_ / model.rho_0_)
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 100, check, reinit)
    x0 = y0 = 0.0
    fluid_width = l * length_factor
    n_row = round(l / dr)
    n_col = round(fluid_width / dr)

    def liquid(X):
        p = LiquidParticles(dim, len(X))
        p.x_vec_ = X
        p.p_ = np.full(len(p), p_0)
        p.rho_ = np.full(len(p), rho_0)
        p.c_ = np.full(len(p), c_0)
        p.mass_ = np.full(len(p), dr ** 2 * rho_0)
        p.gap_ = np.full(len(p), gap)
        p.v_vec_[:, 0] = _channel_ux(X[:, 1] - y0, 0.0, f, nu, l, terms)
        return p

    t = 3
    fl = liquid(rectangle(dr, x0, y0, n_col, n_row))
    inlet = liquid(rectangle(dr, x0 - t * dr, y0, t, n_row))
    outlet = liquid(rectangle(dr, x0 + fluid_width, y0, t, n_row))
    n_more = 3
    ww = round(fluid_width / dr) + 2 * (t + n_more)
    C = _lib.WALL_COMPULSIVE
    groups = [
        _walls(C, rectangle(dr, x0 - (t + n_more) * dr, y0 - t * dr, ww, t), [0.0, 1.0], gap),
        _walls(C, rectangle(dr, x0 - (t + n_more) * dr, y0 + l, ww, t), [0.0, -1.0], gap),
    ]
    walls = WallParticles.concat(groups)
    planes = InletOutletPlanes(0, x0, t * dr, x0 + fluid_width, x0 + fluid_width + t * dr)
    profile = BufferProfile(_lib.PROFILE_CHANNEL_2D, f, nu, l, (x0, y0, 0.0), terms, outlet_profile=True)
    return Scenario("poiseuille_2d", dim, sk, model, td, fl, [walls], None, (inlet, outlet), planes, profile)


def cylinder_ring(dx, x_0, y_0, z_0, inner_r, outer_r, length, rho=None):
    """createCylinderRingParticles (PoiseuilleFlow3D.jl:60-93): rings of
    particles along x; returns positions, masses (when rho is given) and gaps."""
    len_num = round(length / dx)
    len_dx = length / len_num
    r_num = round((outer_r - inner_r) / dx)
    r_dx = (outer_r - inner_r) / r_num
    X, M, G = [], [], []
    for i_len in range(1, len_num + 1):
        x_here = x_0 + (i_len - 0.5) * len_dx
        for i_r in range(1, r_num + 1):
            r_here = inner_r + (i_r - 0.5) * r_dx
            n_theta = round(2 * math.pi * r_here / r_dx)
            if n_theta == 0:
                n_theta = 1
            d_theta = 2 * math.pi / n_theta
            dr_here = 2 * math.pi * r_here / n_theta
            th = np.arange(1, n_theta + 1) * d_theta
            blk = np.empty((n_theta, 3))
            blk[:, 0] = x_here
            blk[:, 1] = y_0 + r_here * np.cos(th)
            blk[:, 2] = z_0 + r_here * np.sin(th)
            X.append(blk)
            M.append(np.full(n_theta, len_dx * dr_here * r_dx * rho if rho is not None else 0.0))
            G.append(np.full(n_theta, (len_dx + dr_here + r_dx) / 3))
    return np.concatenate(X), np.concatenate(M), np.concatenate(G)


def poiseuille_3d(cells=40, l=1e-2, nu=1e-6, rho_0=1000.0, f=8e-5, kernel=WendlandC2, radius_factor=3.0,
                  length_factor=2, total_step=1000, reinit=5, check=5):
    """3D Poiseuille pipe with inlet and outlet buffers
    (PoiseuilleFlow3D.jl:11-158): radius l / 2, dr = l / ``cells`` (the
    example has 40), ``length_factor * l`` long, three buffer rings at each
    end and three compulsive-wall rings around a pipe 12 dr longer.  The
    fluid starts on the steady profile; the outlet keeps its axial velocity
    (only its cross components are zeroed after each step)."""
    dim = 3
    r0 = l / 2
    dr = l / cells
    R = radius_factor * dr
    sk = SmoothKernel(R, dim, kernel if not isinstance(kernel, str) else KERNELS[kernel])
    ux_max = r0 ** 2 / 4 * f / nu
    c_0 = 10 * ux_max
    p_0 = 0.05 * rho_0 * c_0 ** 2
    mu_0 = nu * rho_0
    gamma = 7
    model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, [f, 0.0, 0.0])
    dt = 0.1 * sk.influence_radius_ / c_0
    dt = viscous_dt(dt, sk.h_, model.mu_0_ / model.rho_0_)
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, total_step, 10, check, reinit)
    x0 = y0 = z0 = 0.0
    pipe_len = length_factor * l

    def liquid(X, M, G):
        p = LiquidParticles(dim, len(X))
        p.x_vec_ = X
        p.mass_ = M
        p.gap_ = G
        p.p_ = np.full(len(p), p_0)
        p.rho_ = np.full(len(p), rho_0)
        p.c_ = np.full(len(p), c_0)
        p.v_vec_[:, 0] = f / 4 / nu * (r0 ** 2 - X[:, 1] ** 2 - X[:, 2] ** 2)
        return p

    t = 3
    fl = liquid(*cylinder_ring(dr, x0, y0, z0, 0.0, r0, pipe_len, rho_0))
    inlet = liquid(*cylinder_ring(dr, x0 - t * dr, y0, z0, 0.0, r0, t * dr, rho_0))
    outlet = liquid(*cylinder_ring(dr, x0 + pipe_len, y0, z0, 0.0, r0, t * dr, rho_0))
    WX, _, WG = cylinder_ring(dr, x0 - 2 * t * dr, y0, z0, r0, r0 + t * dr, pipe_len + 4 * t * dr)
    nrm = np.zeros_like(WX)
    rr = np.sqrt(WX[:, 1] ** 2 + WX[:, 2] ** 2)
    nrm[:, 1] = -WX[:, 1] / rr
    nrm[:, 2] = -WX[:, 2] / rr
    walls = WallParticles(_lib.WALL_COMPULSIVE, dim, len(WX))
    walls.x_vec_ = WX
    walls.normal_vec_ = nrm
    walls.gap_ = WG
    planes = InletOutletPlanes(0, x0, t * dr, x0 + pipe_len, x0 + pipe_len + t * dr)
    profile = BufferProfile(_lib.PROFILE_PIPE_3D, f, nu, r0, (x0, y0, z0), 0, outlet_profile=False)
    return Scenario("poiseuille_3d", dim, sk, model, td, fl, [walls], None, (inlet, outlet), planes, profile)


def _log_uniform(rng, lo, hi):
    return float(np.exp(rng.uniform(math.log(lo), math.log(hi))))


def ensemble_2d(n_scenes=256, seed=0, n_min=2000, n_max=60000, total_step=5000):
    """The batched-ensemble workload: alternating lid-driven cavities (lid
    1 m/s, Re uniform in [10, 1000]) and Poiseuille channels (body force
    giving Re uniform in [1, 100]); the fluid count of each scene is
    log-uniform in [n_min, n_max]; CubicSpline or WendlandC2 drawn per scene;
    everything from one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        n = _log_uniform(rng, n_min, n_max)
        kern = CubicSpline if rng.random() < 0.5 else WendlandC2
        if k % 2 == 0:
            side = max(8, round(math.sqrt(n)))
            sc = lid_driven_cavity_2d(dx=1.0 / side, reynolds=float(rng.uniform(10.0, 1000.0)), kernel=kern,
                                      total_step=total_step)
        else:
            rows = max(6, round(math.sqrt(n / 4.0)))
            sc = poiseuille_2d(rows=rows, reynolds=float(rng.uniform(1.0, 100.0)), kernel=kern,
                               total_step=total_step)
        out.append(sc)
    return out


def serving_stream(n_requests=32, seed=0, rate_per_s=2.0, steps=(1000, 10000), scale=1.0):
    """The serving workload: Poisson arrivals (rate_per_s) of requests drawn
    uniformly from three kinds, sizes log-uniform per kind --
    2D dam break 5k-200k fluid particles, 3D Poiseuille pipe 50k-2M, 2D
    thermal cavity 10k-100k -- each running a step count uniform in
    ``steps``.  ``scale`` multiplies every size (for bounded samples).
    Returns [(arrival_s, scenario, nsteps)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        kind = int(rng.integers(3))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        if kind == 0:
            n = _log_uniform(rng, 5e3, 2e5) * scale
            sc = collapse_dry_2d(dx=math.sqrt(2.0 / n), total_step=nsteps)
        elif kind == 1:
            n = _log_uniform(rng, 5e4, 2e6) * scale
            sc = poiseuille_3d(cells=max(6, round((2.0 * n / math.pi) ** (1.0 / 3.0))), total_step=nsteps)
        else:
            n = _log_uniform(rng, 1e4, 1e5) * scale
            sc = thermal_cavity_2d(dx=1.0 / max(8, round(math.sqrt(n))), seed=int(rng.integers(1 << 30)),
                                   total_step=nsteps)
        out.append((t, sc, nsteps))
    return out
