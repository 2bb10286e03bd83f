"""CPU: the oracle's SPH probe interpolation (kernelValueInterpolation,
DataIO/VTPToDataFrame.jl:77-142) against what the mathematics fixes:
constants are reproduced, a linear field is reproduced at lattice-symmetric
points, two-particle weights follow W m / rho, none_value where no particle
is within R."""
import numpy as np
import pytest

from ethersph import _lib, scenarios
from oracle import oracle as O

from helpers import lattice_scene


def _scene_with(sc, rho=None, v=None, p=None):
    if rho is not None:
        sc.fluid.rho_ = np.asarray(rho, dtype=np.float64)
    if v is not None:
        sc.fluid.v_vec_ = np.asarray(v, dtype=np.float64)
    if p is not None:
        sc.fluid.p_ = np.asarray(p, dtype=np.float64)
    return O.from_scenario(sc)


@pytest.mark.parametrize("dim", [2, 3])
def test_constant_and_linear_fields(dim):
    n = 14 if dim == 2 else 8
    dx = 0.01
    sc = lattice_scene(n=n, dx=dx, dim=dim, kernel="CubicSpline")
    X = sc.fluid.x_vec_
    rng = np.random.default_rng(3)
    sc.fluid.mass_ = sc.fluid.mass_ * (1 + 0.2 * rng.random(len(sc.fluid)))
    a = np.array([0.7, -1.3, 2.1][:dim])
    o = _scene_with(sc, rho=np.full(len(X), 1000.0), v=np.tile(a, (len(X), 1)), p=X @ a)
    mid = n * dx / 2
    # probes at lattice sites and at cell centres near the middle
    pts = np.array([[mid + dx / 2] * dim, [mid] * dim, [mid - dx / 2] * dim])
    pts = np.vstack([pts, pts + dx])
    np.testing.assert_allclose(o.probe(pts, _lib.FIELD_V), np.tile(a, (len(pts), 1)), rtol=1e-14)
    np.testing.assert_allclose(o.probe(pts, _lib.FIELD_RHO), 1000.0, rtol=1e-14)


def test_linear_field_needs_symmetric_weights():
    """With equal masses on the lattice, a linear pressure is reproduced at the
    symmetric probe points (the weights are symmetric about them)."""
    dx = 0.01
    sc = lattice_scene(n=16, dx=dx, dim=2, kernel="WendlandC2")
    X = sc.fluid.x_vec_
    a = np.array([3.0, -2.0])
    o = _scene_with(sc, p=X @ a)
    pts = np.array([[0.08, 0.08], [0.085, 0.085], [0.075, 0.085]])
    np.testing.assert_allclose(o.probe(pts, _lib.FIELD_P), pts @ a, rtol=1e-12)


def test_two_particle_weights_and_none_value():
    sc = lattice_scene(n=2, dx=0.01, dim=2, kernel="CubicSpline")
    sc.fluid = sc.fluid.subset([0, 1])
    sc.fluid.x_vec_ = np.array([[0.0, 0.0], [0.01, 0.0]])
    sc.fluid.mass_ = np.array([0.1, 0.3])
    sc.fluid.rho_ = np.array([1000.0, 1200.0])
    sc.fluid.p_ = np.array([5.0, 11.0])
    o = O.from_scenario(sc)
    R = sc.kernel.influence_radius_
    pts = np.array([[0.005, 0.0], [0.0025, 0.0], [0.5, 0.5]])
    got = o.probe(pts, _lib.FIELD_P, none_value=-7.0)
    for q in range(2):
        d = np.linalg.norm(pts[q] - sc.fluid.x_vec_, axis=1)
        W, _, _ = O.kernel_eval(sc.kernel.kind, 2, R, d)  # pinned by the kernel tests
        w = W * sc.fluid.mass_ / sc.fluid.rho_
        assert got[q] == pytest.approx((w * sc.fluid.p_).sum() / w.sum(), rel=1e-14)
    # at the midpoint both W are equal: the m / rho weights alone decide
    wm = sc.fluid.mass_ / sc.fluid.rho_
    assert got[0] == pytest.approx((wm * sc.fluid.p_).sum() / wm.sum(), rel=1e-14)
    assert got[2] == -7.0
    assert np.linalg.norm(pts[2]) > R


def test_probe_on_a_stepped_scene_is_bounded():
    """After steps the interpolated density stays within the particle range."""
    sc = scenarios.collapse_dry_2d(dx=0.05)
    o = O.from_scenario(sc)
    o.step(20)
    rho = o.get(_lib.FIELD_RHO)
    pts = np.column_stack([np.linspace(0.05, 0.95, 25), np.linspace(0.05, 1.9, 25)])
    got = o.probe(pts, _lib.FIELD_RHO)
    assert np.all(got >= rho.min() * (1 - 1e-12)) and np.all(got <= rho.max() * (1 + 1e-12))
