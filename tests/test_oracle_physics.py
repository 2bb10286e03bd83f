"""Pins the oracle's step (continuity, EOS, forces, integration, Shepard,
check) with known answers that a plausible transcription mistake would break:
the reference's own tests, closed forms of special cases, SPH consistency of
each operator on a lattice (sign and factor errors fail these), and
conservation laws of the pair terms."""
import json
import math
import os

import numpy as np
import pytest

from ethersph import _lib
from ethersph.models import CommonWeaklyCompressibleLiquidModel, ThermalWeaklyCompressibleLiquidModel
from oracle import oracle as O

from helpers import AABBCheck, interior, lattice_scene, random_blob, rel_l2, wall_plane

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))


def test_equation_model_reference_test():
    g = GOLD["equation_model"]
    m = CommonWeaklyCompressibleLiquidModel(g["rho_0"], g["c_0"], g["p_0"], g["gamma"], g["mu_0"],
                                            g["body_force_vec"])
    assert math.isclose(m.nu_0_, g["expected_nu_0"])
    t = ThermalWeaklyCompressibleLiquidModel(g["rho_0"], g["c_0"], g["p_0"], g["gamma"], g["mu_0"],
                                             g["body_force_vec"], g["t_0"], g["kappa_0"], g["cp"], g["beta"])
    assert math.isclose(t.b_, g["expected_b"])


@pytest.mark.parametrize("gamma", [1, 7])
def test_eos_closed_form(gamma):
    """One isolated particle: the step's EOS gives p = c0^2 rho0/gamma((rho/rho0)^gamma - 1) + p0,
    c = c0 sqrt((rho/rho0)^(gamma-1)); gamma = 1 is the linear EOS p = c0^2 (rho - rho0) + p0."""
    sc = lattice_scene(n=1, gamma=gamma, c_0=30.0, p_0=5.0)
    sc.fluid.rho_[:] = 1013.0
    o = O.from_scenario(sc)
    o.step(1)
    rr = 1013.0 / 1000.0
    p = 30.0 ** 2 * 1000.0 / gamma * (rr ** gamma - 1.0) + 5.0
    c = 30.0 * math.sqrt(rr ** (gamma - 1))
    assert math.isclose(o.get(_lib.FIELD_P)[0], p, rel_tol=1e-13)
    assert math.isclose(o.get(_lib.FIELD_C)[0], c, rel_tol=1e-13)
    if gamma == 1:
        assert math.isclose(o.get(_lib.FIELD_P)[0], 30.0 ** 2 * (1013.0 - 1000.0) + 5.0, rel_tol=1e-12)
        assert o.get(_lib.FIELD_C)[0] == 30.0


def test_free_particle_under_gravity():
    """No neighbours: v_n = n g dt, x_n = x0 + g dt^2 n (n + 1) / 2 (the
    reference updates v first, then x with the new v)."""
    sc = lattice_scene(n=1, g=[0.3, -9.81], dt=1e-3)
    x0 = sc.fluid.x_vec_[0].copy()
    o = O.from_scenario(sc)
    n = 57
    o.step(n)
    g = np.array([0.3, -9.81])
    assert np.allclose(o.get(_lib.FIELD_V)[0], n * g * 1e-3, rtol=1e-12)
    assert np.allclose(o.get(_lib.FIELD_X)[0], x0 + g * 1e-6 * n * (n + 1) / 2, rtol=1e-12, atol=1e-15)


def test_neighbour_345_reference_test():
    """test/Neighbour.jl: x_i = (3, 4), x_j = 0, R = 10 -> r = 5.  Two
    particles; continuity gives drho_i = m_j (v_ij . W'(5) (3,4)/5)."""
    g = GOLD["neighbour_345"]
    pairs = O.neighbor_list(np.array([g["x_i"], g["x_j"]]), g["influence_radius"])
    assert pairs.tolist() == [[0, 1]]
    # r = 5 with h = 5: q = 1, CubicSpline W' = -3 sigma / (4 h)
    w, dw, p = O.kernel_eval(0, 2, g["influence_radius"], [g["expected_r"]])
    sc = lattice_scene(n=2, kernel="CubicSpline", h_factor=1.0, dx=5.0, dt=1e-6)
    assert sc.kernel.influence_radius_ == 10.0 and sc.kernel.h_ == 5.0
    f = sc.fluid.subset([0, 1])
    f.x_vec_ = np.array([g["x_i"], g["x_j"]], dtype=float)
    f.v_vec_ = np.array([[1.0, 2.0], [0.0, 0.0]])
    f.mass_ = np.array([2.0, 3.0])
    o = O.OracleScene(sc.desc(), f, [])
    o.step(1)
    grad = dw[0] * np.array([3.0, 4.0]) / 5.0
    drho_i = 3.0 * np.dot([1.0, 2.0], grad)
    drho_j = 2.0 * np.dot([1.0, 2.0], grad)
    rho = o.get(_lib.FIELD_RHO)
    assert math.isclose(rho[0], 1000.0 + drho_i * 1e-6, rel_tol=1e-14)
    assert math.isclose(rho[1], 1000.0 + drho_j * 1e-6, rel_tol=1e-14)


def _one_step_rates(sc):
    o = O.from_scenario(sc)
    rho0 = sc.fluid.rho_.copy()
    v0 = sc.fluid.v_vec_.copy()
    o.step(1)
    dt = sc.td.dt_
    return o, (o.get(_lib.FIELD_RHO) - rho0) / dt, (o.get(_lib.FIELD_V) - v0) / dt


@pytest.mark.parametrize("kernel", ["CubicSpline", "WendlandC2", "WendlandC4"])
def test_continuity_is_minus_rho_div_v(kernel):
    sc = lattice_scene(n=30, kernel=kernel)
    X = sc.fluid.x_vec_
    sc.fluid.v_vec_ = np.stack([0.3 * X[:, 0], -0.1 * X[:, 1]], 1)
    _, drho, _ = _one_step_rates(sc)
    m = interior(X, 0.0, 0.3, 0.05)
    assert np.allclose(drho[m], -1000.0 * (0.3 - 0.1), rtol=2e-2)


def test_pressure_force_is_minus_grad_p_over_rho():
    """Linear pressure field (linear EOS, zero velocity): dv = -grad p / rho."""
    sc = lattice_scene(n=30, kernel="WendlandC4", c_0=10.0)
    X = sc.fluid.x_vec_
    gp = 2000.0  # Pa / m
    sc.fluid.rho_ = 1000.0 + (1000.0 + gp * X[:, 1]) / 100.0  # p = c0^2 (rho - rho0)
    _, _, dv = _one_step_rates(sc)
    m = interior(X, 0.0, 0.3, 0.05)
    rho = sc.fluid.rho_[m]
    # lattice discretisation + the 1% tensile term: a few percent
    assert np.allclose(dv[m, 1], -gp / rho, rtol=6e-2)
    assert np.abs(dv[m, 0]).max() < 1e-2 * gp / 1000.0


def test_viscosity_is_nu_laplacian():
    """Shear profile v_x = a y^2: the viscous term gives nu * 2a in the interior."""
    a = 5.0
    sc = lattice_scene(n=34, kernel="WendlandC2", mu_0=0.5)
    X = sc.fluid.x_vec_
    sc.fluid.v_vec_ = np.stack([a * X[:, 1] ** 2, 0 * X[:, 1]], 1)
    _, _, dv = _one_step_rates(sc)
    m = interior(X, 0.0, 0.34, 0.06)
    nu = 0.5 / 1000.0
    assert np.allclose(dv[m, 0], 2 * a * nu, rtol=5e-2)


def test_thermal_conduction_is_diffusion():
    """T = y^2, uniform kappa, cp, rho: dT/dt = kappa lap(T) / (rho cp) = 2 kappa / (rho cp)."""
    sc = lattice_scene(n=30, kernel="WendlandC2", thermal=True)
    X = sc.fluid.x_vec_
    sc.fluid.t_ = X[:, 1] ** 2
    sc.fluid.kappa_[:] = 3.0
    sc.fluid.cp_[:] = 2.0
    walls = [wall_plane(2, 30, 0.01, kind=_lib.WALL_THERMOSTATIC, t=0.0, kappa=3.0)]
    sc.walls = walls
    o = O.from_scenario(sc)
    T0 = sc.fluid.t_.copy()
    o.step(1)
    dT = (o.get(_lib.FIELD_T) - T0) / sc.td.dt_
    m = interior(X, 0.0, 0.3, 0.05)
    assert np.allclose(dT[m], 2 * 3.0 / (1000.0 * 2.0), rtol=5e-2)


def test_shepard_reproduces_constants_and_two_particle_case():
    sc = lattice_scene(n=12, kernel="CubicSpline", reinit=1)
    o = O.from_scenario(sc)
    o.step(1)
    assert np.allclose(o.get(_lib.FIELD_RHO), 1000.0, rtol=1e-13)
    # two isolated particles: rho_1 <- (W0 m1 + W12 m2) / (W0 m1/rho1 + W12 m2/rho2)
    sc = lattice_scene(n=2, kernel="CubicSpline", reinit=1)
    f = sc.fluid.subset([0, 1])
    f.x_vec_ = np.array([[0.0, 0.0], [0.017, 0.0]])
    f.rho_ = np.array([1000.0, 1100.0])
    f.mass_ = np.array([0.1, 0.13])
    o = O.OracleScene(sc.desc(), f, [])
    o.step(1)
    w, _, p = O.kernel_eval(0, 2, sc.kernel.influence_radius_, [0.017])
    W0, W = p["kernel_0"], w[0]
    r1 = (W0 * 0.1 + W * 0.13) / (W0 * 0.1 / 1000.0 + W * 0.13 / 1100.0)
    r2 = (W0 * 0.13 + W * 0.1) / (W0 * 0.13 / 1100.0 + W * 0.1 / 1000.0)
    assert np.allclose(o.get(_lib.FIELD_RHO), [r1, r2], rtol=1e-13)


def test_momentum_and_heat_conservation():
    """Fluid-fluid pair terms are antisymmetric: sum m dv = 0 and
    sum m cp dT = 0 for an isolated blob without body force."""
    sc = random_blob(n=400, thermal=True, kernel="CubicSpline")
    f = sc.fluid
    P0 = (f.mass_[:, None] * f.v_vec_).sum(0)
    H0 = (f.mass_ * f.cp_ * f.t_).sum()
    d = sc.desc()
    d.solver = _lib.SOLVER_FLUID_THERMOSTATIC_WALL  # thermal solver, no wall particles
    o = O.OracleScene(d, f, [])
    o.step(1)
    dt = sc.td.dt_
    v1 = o.get(_lib.FIELD_V)
    P1 = (f.mass_[:, None] * v1).sum(0)
    scale = (f.mass_[:, None] * np.abs(v1 - f.v_vec_)).sum() + 1e-300
    assert np.abs(P1 - P0).max() < 1e-12 * scale
    H1 = (f.mass_ * f.cp_ * o.get(_lib.FIELD_T)).sum()
    assert abs(H1 - H0) < 1e-12 * H0


def test_wall_force_single_pair_closed_form():
    """Fluid particle at distance d = 0.5 gap above one wall particle, v = 0:
    dv = (0.01 c^2 / h) |1 - q| / sqrt(q) * n with q = 0.5, P(xi = 0) = 1."""
    dx = 0.01
    sc = lattice_scene(n=1, kernel="WendlandC2", c_0=10.0, dt=1e-6)
    sc.fluid.x_vec_ = np.array([[0.0, 0.5 * dx]])
    w = wall_plane(2, 1, dx, layers=1)
    w.x_vec_ = np.array([[0.0, 0.0]])
    w.normal_vec_ = np.array([[0.0, 1.0]])
    w.gap_ = np.array([dx])
    sc.walls = [w]
    o = O.from_scenario(sc)
    o.step(1)
    h = sc.kernel.h_
    q = 0.5
    R = 0.01 * 10.0 ** 2 / h * abs(1 - q) / math.sqrt(q)
    # the wall viscosity term vanishes (v = 0) and FluidWallSolve applies no wall pressure
    dv = o.get(_lib.FIELD_V)[0] / 1e-6
    assert math.isclose(dv[1], R, rel_tol=1e-12) and abs(dv[0]) < 1e-12 * R
    # beyond the gap (q > 1) the force is zero
    sc.fluid.x_vec_ = np.array([[0.0, 1.2 * dx]])
    o = O.from_scenario(sc)
    o.step(1)
    assert np.all(o.get(_lib.FIELD_V)[0] == 0.0)


def test_check_deletes_stably():
    sc = lattice_scene(n=10, check_interval=2, check=AABBCheck(hi=[0.07, 1.0]))
    sc.fluid.v_vec_[:, 0] = 0.0
    o = O.from_scenario(sc)
    o.step(1)
    assert o.fluid_count() == 100  # not a check step
    o.step(1)
    ids = o.get(_lib.FIELD_ID).astype(int)
    X = sc.fluid.x_vec_
    expect = np.nonzero(X[:, 0] <= 0.07)[0]
    assert np.array_equal(ids, expect)


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbor_list_cell_equals_brute_and_numpy_rule(dim):
    rng = np.random.default_rng(dim)
    n = 1500 if dim == 2 else 2500
    X = rng.uniform(0, 1, (n, dim))
    R = 0.05 if dim == 2 else 0.12
    a = O.neighbor_list(X, R, brute=True)
    b = O.neighbor_list(X, R)
    assert np.array_equal(a, b)
    # independent restatement of the rule in numpy: each op rounded (no FMA)
    d = X[:, None, :] - X[None, :, :]
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    if dim == 3:
        d2 = d2 + d[..., 2] * d[..., 2]
    i, j = np.nonzero(np.triu(d2 <= R * R, 1))
    assert np.array_equal(a, np.stack([i, j], 1))
    Y = rng.uniform(0, 1, (700, dim))
    c = O.neighbor_list(X, R, Y, brute=True)
    e = O.neighbor_list(X, R, Y)
    assert np.array_equal(c, e)


def test_exact_cutoff_lattice_pairs():
    """2D dam-break lattice: offsets (3,0) and (0,3) sit at R = 3 dx; inclusion
    follows the binary64 rule pair by pair (some rounded in, some out)."""
    from ethersph import scenarios

    sc = scenarios.collapse_dry_2d()
    X = sc.fluid.x_vec_
    R = sc.kernel.influence_radius_
    pairs = O.neighbor_list(X, R)
    d = X[pairs[:, 0]] - X[pairs[:, 1]]
    assert np.all(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] <= R * R)
    on_cut = np.isclose(np.hypot(d[:, 0], d[:, 1]), R, rtol=1e-12)
    assert 0 < on_cut.sum() < 9800  # some, not all, of the 3-dx lattice offsets are in


def test_check_planes_delete_like_deleteat():
    """A general check (box plus tilted half-spaces) on free particles:
    the survivors after N steps are those a replay of check + deleteat! on
    Python lists keeps, in order (FluidWallSolve.jl:73-81)."""
    rng = np.random.default_rng(11)
    n = 60
    sc = lattice_scene(n=1, dx=0.001, check_interval=3,
                       check=AABBCheck(lo=[-0.5, -0.5], hi=[0.5, 0.4],
                                       planes=[((0.6, 0.8), 0.3), ((-1.0, 1.0), 0.35)]),
                       dt=1e-3)
    f = sc.fluid.subset(np.zeros(n, dtype=int))
    # isolated particles: one per row 0.011 apart (rows farther apart than R),
    # moving along x at constant speed
    f.x_vec_ = np.column_stack([rng.uniform(-0.3, 0.3, n), np.arange(n) * 0.011 - 0.33])
    f.v_vec_ = np.column_stack([rng.uniform(-40.0, 40.0, n), np.zeros(n)])
    sc.fluid = f
    R = sc.kernel.influence_radius_
    assert R < 0.011
    o = O.from_scenario(sc)
    steps = 30
    o.step(steps)
    # replay
    xs = [np.array(p, dtype=np.float64) for p in f.x_vec_]
    vs = [np.array(v, dtype=np.float64) for v in f.v_vec_]
    ids = list(range(n))
    dt = sc.td.dt_
    for step in range(1, steps + 1):
        xs = [x + v * dt - 0.0 * (dt * dt) / 2.0 for x, v in zip(xs, vs)]
        if step % 3 == 0:
            keep = []
            for k, x in enumerate(xs):
                out = x[0] > 0.5 or x[1] > 0.4 or x[0] < -0.5 or x[1] < -0.5
                out = out or (0.6 * x[0] + 0.8 * x[1]) > 0.3 or (-1.0 * x[0] + 1.0 * x[1]) > 0.35
                if not out:
                    keep.append(k)
            xs = [xs[k] for k in keep]
            vs = [vs[k] for k in keep]
            ids = [ids[k] for k in keep]
    assert 0 < len(ids) < n
    assert np.array_equal(o.get(_lib.FIELD_ID).astype(int), ids)
    np.testing.assert_array_equal(o.get(_lib.FIELD_X), np.array(xs))
