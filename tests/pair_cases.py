"""Single-pair scenes with closed-form answers for the fluid-wall terms,
the relative velocity of static-velocity walls, Boussinesq buoyancy and the
coincident-particle gradient guard.

Each case is one fluid particle (or two) and at most one wall particle, one
step, no body force unless the case is about it.  The expected values are
evaluated here from the reference's formulas with numpy, independently of
the oracle and of the device code:

- FW pressure  PressureForce.jl:23-35   dv_i = -m_i P gradW,
  P = p_i / rho_i^2 + p_0 / rho_0^2, P += |P| 0.01 W(r) / W(gap_i)
- FW viscosity ViscosityForce.jl:27-46  dv_i = m_i f (v_ij - (v_ij . n) n),
  f = 8 mu_i r W' / (2 rho_i)^2 / (r^2 + 0.01 h^2), mu_i = nu_0 rho_i
- FW conduction ThermalConduction.jl:24-36  dT_i = m_i dq,
  dq = (kappa_i + kappa_w) r W' / cp_model / rho_i / rho_0 / (r^2 + 0.01 h^2) (T_i - T_w)
- buoyancy     FluidThermostaticWallSolve.jl:106-110,
  FluidCompulsiveThermostaticWallSolve.jl:134-138:
  T^{n+1} = T^n + dT dt first, then v += -g beta (T^{n+1} - t_0) dt
- wall force   WallForce.jl:8-30 with v_ij = v_i - v_wall for
  StaticVelocityWallParticle (Neighbour.jl:24-29) and v_ij = v_i for a
  fixed wall (Neighbour.jl:17-22)
- gradW guard  Neighbour.jl:54-59: normalize(0) is NaN and is zeroed, so a
  coincident pair exchanges nothing, while W(0) still enters Shepard.

Cases place the fluid particle more than one wall gap from the wall particle
(q > 1) unless the case is about the wall force, so WallForce.jl returns
early; zero velocities or p = p_0 = 0 switch off the terms a case is not
about.  A dropped term, a wrong sign, m_j for m_i, the particle's cp for the
model's, T^n for T^{n+1} or v_i for v_i - v_wall all change a checked value.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ethersph import _lib
from ethersph.models import CommonWeaklyCompressibleLiquidModel, ThermalWeaklyCompressibleLiquidModel
from ethersph.particles import LiquidParticles, WallParticles
from ethersph.scenarios import Scenario
from ethersph.sph_kernel import KERNELS, SmoothKernel
from ethersph.time_discretization import DensityReinitializedFixedTimeIntervalForwardEuler

DX = 0.01
H = 1.5 * DX          # WendlandC2 and CubicSpline: R = 2 h = 3 dx
DT = 1e-5


def wendland_c2_2d(r, h=H):
    """W and W' of WendlandC2 in 2D (sigma = 7 / (4 pi h^2)), WendlandC2.jl:38-63."""
    s = 7.0 / (4.0 * math.pi * h * h)
    q = r / h
    if q >= 2.0:
        return 0.0, 0.0
    return s * (2 - q) ** 4 * (1 + 2 * q) / 16.0, -s / h * 5.0 / 8.0 * q * (2 - q) ** 3


def cubic_spline_2d(r, h=H):
    """W and W' of CubicSpline in 2D (sigma = 10 / (7 pi h^2)), CubicSpline.jl:39-73."""
    s = 10.0 / (7.0 * math.pi * h * h)
    q = r / h
    if q < 1.0:
        return s / 4.0 * (3 * q * q * (q - 2) + 4), s / h * 0.75 * q * (3 * q - 4)
    if q < 2.0:
        return s / 4.0 * (2 - q) ** 3, -s / h * 0.75 * (2 - q) ** 2
    return 0.0, 0.0


@dataclass
class Case:
    name: str
    sc: Scenario
    expect: dict = field(default_factory=dict)   # field id -> expected array after `steps` steps
    rtol: float = 1e-12
    atol: dict = field(default_factory=dict)     # field id -> absolute floor
    steps: int = 1


def _fluid(x, v, rho, mass, thermal=False, t=0.0, kappa=0.0, cp=0.0, c=10.0, p=0.0):
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    f = LiquidParticles(2, x.shape[0], thermal)
    f.x_vec_ = x.copy()
    f.v_vec_ = np.atleast_2d(np.asarray(v, dtype=np.float64)).copy()
    f.rho_ = np.full(len(f), float(rho)) if np.isscalar(rho) else np.asarray(rho, dtype=np.float64)
    f.mass_ = np.full(len(f), float(mass)) if np.isscalar(mass) else np.asarray(mass, dtype=np.float64)
    f.gap_ = np.full(len(f), DX)
    f.c_ = np.full(len(f), c)
    f.p_ = np.full(len(f), p)
    if thermal:
        f.t_ = np.full(len(f), t)
        f.kappa_ = np.full(len(f), kappa)
        f.cp_ = np.full(len(f), cp)
    return f


def _wall(kind, x, normal, v=None, t=0.0, kappa=0.0, gap=DX):
    w = WallParticles(kind, 2, 1)
    w.x_vec_ = np.array([x], dtype=np.float64)
    w.normal_vec_ = np.array([normal], dtype=np.float64)
    w.gap_ = np.array([gap])
    if v is not None:
        w.v_vec_ = np.array([v], dtype=np.float64)
    w.t_ = np.array([t])
    w.kappa_ = np.array([kappa])
    return w


def _scenario(name, fluid, walls, model, kernel="WendlandC2", reinit=0):
    K = KERNELS[kernel]
    sk = SmoothKernel(K.radius_ratio * H, 2, K)
    td = DensityReinitializedFixedTimeIntervalForwardEuler(DT, 1, 1, 0, reinit)
    return Scenario(name, 2, sk, model, td, fluid, walls, None)


def _expect_v_x(x0, v0, dv, g=(0.0, 0.0)):
    """updateVelocity! then updatePosition! (UpdateVelocity.jl:8-23,
    UpdatePosition.jl:8-14): v1 = v0 + g dt + dv dt; x1 = x0 + v1 dt - dv dt^2 / 2."""
    v1 = np.asarray(v0) + np.asarray(g) * DT + np.asarray(dv) * DT
    x1 = np.asarray(x0) + v1 * DT - np.asarray(dv) * DT * DT / 2.0
    return v1, x1


def fw_pressure(tilted: bool) -> Case:
    """FW pressure with a static-velocity wall at rest (FluidVelocitySolve
    applies it): rho = 1010 gives p = c0^2 rho0 (rho/rho0 - 1) + p0 (gamma 1)."""
    c0, rho0, p0, rho, m = 10.0, 1000.0, 5.0, 1010.0, 0.11
    xw = np.array([0.0, 0.0])
    xi = np.array([0.7 * DX, 1.3 * DX]) if tilted else np.array([0.0, 1.5 * DX])
    f = _fluid(xi, [0.0, 0.0], rho, m)
    w = _wall(_lib.WALL_STATIC_VELOCITY, xw, [0.0, 1.0], v=[0.0, 0.0])
    model = CommonWeaklyCompressibleLiquidModel(rho0, c0, p0, 1, 1e-3, [0.0, 0.0])
    sc = _scenario("fw_pressure", f, [w], model)
    assert sc.solver == _lib.SOLVER_FLUID_VELOCITY
    p = c0 * c0 * rho0 * (rho / rho0 - 1.0) + p0
    rv = xi - xw
    r = float(np.linalg.norm(rv))
    W, dW = wendland_c2_2d(r)
    Wg, _ = wendland_c2_2d(DX)
    P = p / rho ** 2 + p0 / rho0 ** 2
    P += abs(P) * 0.01 * W / Wg
    dv = -m * P * dW * rv / r
    v1, x1 = _expect_v_x(xi, [0.0, 0.0], dv)
    return Case(f"fw_pressure{'_tilted' if tilted else ''}", sc,
                {_lib.FIELD_V: v1[None], _lib.FIELD_X: x1[None], _lib.FIELD_P: np.array([p])},
                atol={_lib.FIELD_V: 1e-12 * np.abs(v1).max()})


def fw_pressure_absent_in_fluid_wall_solve() -> Case:
    """The same particle next to a compulsive wall under FluidWallSolve,
    which applies no FW pressure (FluidWallSolve.jl:52-62): v stays 0."""
    c = fw_pressure(False)
    f = c.sc.fluid
    w = _wall(_lib.WALL_COMPULSIVE, [0.0, 0.0], [0.0, 1.0])
    sc = _scenario("fw_no_pressure", f, [w], c.sc.model)
    assert sc.solver == _lib.SOLVER_FLUID_WALL
    return Case("fw_pressure_absent", sc, {_lib.FIELD_V: np.zeros((1, 2)), _lib.FIELD_X: f.x_vec_.copy()},
                atol={_lib.FIELD_V: 0.0})


def fw_viscosity(normal, v, kind=_lib.WALL_COMPULSIVE, v_wall=None, name="fw_viscosity") -> Case:
    """FW viscosity alone (FluidWallSolve: no FW pressure; q > 1: no wall
    force); only the tangential part of v_ij = v_i (- v_wall) acts."""
    rho0, c0, mu0, rho, m = 1000.0, 10.0, 0.7, 1003.0, 0.13
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    tng = np.array([n[1], -n[0]])
    xw = np.array([0.01, 0.02])
    xi = xw + 1.2 * DX * n + 0.4 * DX * tng
    v = np.asarray(v, dtype=np.float64)
    f = _fluid(xi, v, rho, m)
    w = _wall(kind, xw, n, v=v_wall)
    # p_0 = 0 and rho -> p = c0^2 (rho - rho0) != 0: under FluidVelocitySolve
    # the wall pressure would act, so the static-velocity variant sets rho = rho0
    if kind == _lib.WALL_STATIC_VELOCITY:
        rho = rho0
        f.rho_[:] = rho0
    model = CommonWeaklyCompressibleLiquidModel(rho0, c0, 0.0, 1, mu0, [0.0, 0.0])
    sc = _scenario(name, f, [w], model)
    rv = xi - xw
    r = float(np.linalg.norm(rv))
    _, dW = wendland_c2_2d(r)
    nu0 = mu0 / rho0
    fv = 8.0 * (nu0 * rho) * r * dW / (2.0 * rho) ** 2 / (r * r + 0.01 * H * H)
    vij = v - (np.asarray(v_wall, dtype=np.float64) if kind == _lib.WALL_STATIC_VELOCITY else 0.0)
    dv = m * fv * (vij - np.dot(vij, n) * n)
    v1, x1 = _expect_v_x(xi, v, dv)
    scale = max(abs(m * fv) * np.linalg.norm(vij), 1e-300)
    return Case(name, sc, {_lib.FIELD_V: v1[None], _lib.FIELD_X: x1[None]},
                atol={_lib.FIELD_V: 1e-12 * scale * DT + 1e-15 * np.abs(v1).max()})


def fw_thermostat_and_buoyancy(compulsive_group: bool) -> Case:
    """FW-thermostat conduction and Boussinesq buoyancy with T^{n+1}.
    rho = rho0 and p0 = 0 give p = 0 (no pressure), v = 0 (no viscosity),
    q > 1 (no wall force), T^n = t_0 (buoyancy with T^n would be zero)."""
    rho0, c0, cp_model, beta, t0 = 1.0, 4.0, 1.7, 0.3, 0.8
    m, kap_i, cp_i, T_i = 2.3e-4, 2.0, 3.1, t0
    T_w, kap_w = 0.2, 1.5
    g = np.array([0.0, -1.0])
    xw = np.array([0.0, 0.0])
    xi = np.array([0.2 * DX, 1.4 * DX])
    f = _fluid(xi, [0.0, 0.0], rho0, m, thermal=True, t=T_i, kappa=kap_i, cp=cp_i, c=c0)
    wt = _wall(_lib.WALL_THERMOSTATIC, xw, [0.0, 1.0], t=T_w, kappa=kap_w)
    model = ThermalWeaklyCompressibleLiquidModel(rho0, c0, 0.0, 7, 1e-3, list(g), t0, 1.0, cp_model, beta)
    walls = [wt]
    if compulsive_group:
        far = _wall(_lib.WALL_COMPULSIVE, [5.0, 5.0], [0.0, 1.0])
        walls = [far, wt]
    sc = _scenario("fw_thermostat", f, walls, model)
    assert sc.solver == (_lib.SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL if compulsive_group
                         else _lib.SOLVER_FLUID_THERMOSTATIC_WALL)
    rv = xi - xw
    r = float(np.linalg.norm(rv))
    _, dW = wendland_c2_2d(r)
    dq = 1.0 / cp_model * (kap_i + kap_w) * r * dW / rho0 / rho0 / (r * r + 0.01 * H * H) * (T_i - T_w)
    T1 = T_i + dq * m * DT
    bf = -g * beta * (T1 - t0)
    v1, x1 = _expect_v_x(xi, [0.0, 0.0], [0.0, 0.0], g=bf)
    return Case("fw_thermostat" + ("_compulsive" if compulsive_group else ""), sc,
                {_lib.FIELD_T: np.array([T1]), _lib.FIELD_V: v1[None], _lib.FIELD_X: x1[None]},
                atol={_lib.FIELD_V: 1e-12 * abs(v1[1])})


def static_velocity_wall_force(moving: bool) -> Case:
    """Wall force with q = 0.5 and xi = 0 (P(xi) = 1); the fluid particle at
    rest, the wall moving along its normal at 0.4 (static-velocity wall):
    v_ij . n = -0.4 < 0 so beta = 1 and R = (0.01 c^2 + c |v_n|) / h |1 - q| / sqrt(q).
    The fixed-wall variant has v_ij = v_i = 0: R = 0.01 c^2 / h |1 - q| / sqrt(q).
    rho = rho0, p0 = 0 (no pressure); v_ij normal (no viscosity)."""
    rho0, c0, m = 1000.0, 10.0, 0.1
    xi = np.array([0.0, 0.5 * DX])
    f = _fluid(xi, [0.0, 0.0], rho0, m, c=c0)
    kind = _lib.WALL_STATIC_VELOCITY if moving else _lib.WALL_COMPULSIVE
    w = _wall(kind, [0.0, 0.0], [0.0, 1.0], v=[0.0, 0.4] if moving else None)
    model = CommonWeaklyCompressibleLiquidModel(rho0, c0, 0.0, 1, 1e-3, [0.0, 0.0])
    sc = _scenario("wall_force", f, [w], model)
    q = 0.5
    vn = -0.4 if moving else 0.0
    beta = 0.0 if vn > 0 else 1.0
    R = (0.01 * c0 * c0 + beta * c0 * abs(vn)) / H * abs(1.0 - q) / math.sqrt(q)
    dv = np.array([0.0, R])
    v1, x1 = _expect_v_x(xi, [0.0, 0.0], dv)
    return Case("static_velocity_wall_force" if moving else "fixed_wall_force", sc,
                {_lib.FIELD_V: v1[None], _lib.FIELD_X: x1[None]}, atol={_lib.FIELD_V: 1e-12 * R * DT})


def coincident_pair(kernel: str, reinit: int) -> Case:
    """Two fluid particles at one position (r = 0): gradW = W'(0) normalize(0)
    is zeroed, so continuity, pressure and viscosity exchange nothing and
    each particle moves with its own velocity; W(0) = sigma still enters
    Shepard: rho_i <- (W0 m_i + W0 m_j) / (W0 m_i / rho_i + W0 m_j / rho_j)."""
    rho = np.array([1000.0, 1040.0])
    m = np.array([0.1, 0.13])
    x = np.array([[0.3, 0.2], [0.3, 0.2]])
    v = np.array([[0.5, -0.25], [-0.75, 1.0]])
    f = _fluid(x, v, rho, m)
    model = CommonWeaklyCompressibleLiquidModel(1000.0, 10.0, 0.0, 1, 1e-3, [0.0, 0.0])
    sc = _scenario(f"coincident_{kernel}", f, [], model, kernel=kernel, reinit=reinit)
    v1, x1 = _expect_v_x(x, v, np.zeros_like(v))
    expect = {_lib.FIELD_V: v1, _lib.FIELD_X: x1}
    if reinit == 1:
        # W0 cancels: rho_i <- (m_i + m_j) / (m_i / rho_i + m_j / rho_j)
        s = (m[0] + m[1]) / (m[0] / rho[0] + m[1] / rho[1])
        expect[_lib.FIELD_RHO] = np.array([s, s])
    else:
        expect[_lib.FIELD_RHO] = rho.copy()
    return Case(f"coincident_{kernel}_reinit{reinit}", sc, expect, atol={_lib.FIELD_V: 0.0})


def all_cases() -> list[Case]:
    return [
        fw_pressure(False),
        fw_pressure(True),
        fw_pressure_absent_in_fluid_wall_solve(),
        fw_viscosity([0.0, 1.0], [0.3, -0.2]),
        fw_viscosity([0.6, 0.8], [0.3, 0.4], name="fw_viscosity_normal_only"),
        fw_viscosity([0.6, 0.8], [0.8, -0.6], name="fw_viscosity_tangential"),
        fw_viscosity([0.0, 1.0], [0.0, 0.0], kind=_lib.WALL_STATIC_VELOCITY, v_wall=[0.25, 0.0],
                     name="static_velocity_viscosity"),
        fw_thermostat_and_buoyancy(False),
        fw_thermostat_and_buoyancy(True),
        static_velocity_wall_force(True),
        static_velocity_wall_force(False),
        coincident_pair("CubicSpline", 0),
        coincident_pair("WendlandC2", 1),
    ]


def check_fields(get, case: Case, rtol: float | None = None) -> None:
    """Compares get(field) with the case's expected values: |a - e| <=
    rtol |e| + atol per element."""
    rt = case.rtol if rtol is None else rtol
    for fid, e in case.expect.items():
        a = np.asarray(get(fid), dtype=np.float64).reshape(np.shape(e))
        assert np.all(np.isfinite(a)), f"{case.name}: field {fid} not finite: {a}"
        tol = rt * np.abs(e) + case.atol.get(fid, 0.0)
        bad = np.abs(a - e) > tol
        assert not bad.any(), f"{case.name}: field {fid}: got {a.tolist()}, expected {np.asarray(e).tolist()}"
