"""Small scene builders shared by the tests."""
import numpy as np

from ethersph import _lib
from ethersph.models import CommonWeaklyCompressibleLiquidModel, ThermalWeaklyCompressibleLiquidModel
from ethersph.particles import LiquidParticles, WallParticles
from ethersph.scenarios import Scenario, box, rectangle
from ethersph.solve import AABBCheck
from ethersph.sph_kernel import KERNELS, SmoothKernel
from ethersph.time_discretization import DensityReinitializedFixedTimeIntervalForwardEuler


def lattice_scene(n=24, dx=0.01, dim=2, kernel="WendlandC2", h_factor=1.5, rho_0=1000.0, c_0=10.0, gamma=1,
                  mu_0=1e-3, g=None, dt=1e-5, reinit=0, check_interval=0, thermal=False, walls=None,
                  check=None, p_0=0.0):
    K = KERNELS[kernel]
    sk = SmoothKernel(K.radius_ratio * h_factor * dx, dim, K)
    g = [0.0] * dim if g is None else g
    if thermal:
        model = ThermalWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, g, 0.5, 1.0, 1.0, 0.0)
    else:
        model = CommonWeaklyCompressibleLiquidModel(rho_0, c_0, p_0, gamma, mu_0, g)
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, 1, 1, check_interval, reinit)
    X = rectangle(dx, 0.0, 0.0, n, n) if dim == 2 else box(dx, 0.0, 0.0, 0.0, n, n, n)
    f = LiquidParticles(dim, X.shape[0], thermal)
    f.x_vec_ = X
    f.rho_ = np.full(len(f), rho_0)
    f.mass_ = f.rho_ * dx ** dim
    f.gap_ = np.full(len(f), dx)
    f.c_ = np.full(len(f), c_0)
    if thermal:
        f.t_ = np.full(len(f), 0.5)
        f.kappa_ = np.full(len(f), 1.0)
        f.cp_ = np.full(len(f), 1.0)
    return Scenario("lattice", dim, sk, model, td, f, walls or [], check)


def interior(X, lo, hi, margin):
    m = np.ones(X.shape[0], dtype=bool)
    for k in range(X.shape[1]):
        m &= (X[:, k] > lo + margin) & (X[:, k] < hi - margin)
    return m


def random_blob(n=400, dim=2, seed=1, spacing=0.01, thermal=False, kernel="CubicSpline", mu_0=1e-3, gamma=7,
                c_0=20.0, walls=None, dt=1e-5, reinit=0):
    rng = np.random.default_rng(seed)
    side = int(round(n ** (1.0 / dim)))
    sc = lattice_scene(n=side, dx=spacing, dim=dim, kernel=kernel, mu_0=mu_0, gamma=gamma, c_0=c_0,
                       thermal=thermal, walls=walls, dt=dt, reinit=reinit, p_0=0.0)
    f = sc.fluid
    f.x_vec_ = f.x_vec_ + rng.uniform(-0.3, 0.3, f.x_vec_.shape) * spacing
    f.v_vec_ = rng.uniform(-0.5, 0.5, f.v_vec_.shape)
    f.rho_ = 1000.0 * (1.0 + rng.uniform(-0.01, 0.01, len(f)))
    f.mass_ = 1000.0 * spacing ** dim * (1.0 + rng.uniform(-0.05, 0.05, len(f)))
    if thermal:
        f.t_ = rng.uniform(0.0, 1.0, len(f))
        f.kappa_ = rng.uniform(0.5, 2.0, len(f))
        f.cp_ = rng.uniform(0.5, 2.0, len(f))
    return sc


def wall_plane(dim, n_per_side, dx, kind=_lib.WALL_COMPULSIVE, layers=3, v=None, t=None, kappa=None):
    """A flat wall below y = 0 (2D) or z = 0 (3D), normal +y / +z."""
    if dim == 2:
        X = rectangle(dx, -4 * dx, 0.0 - layers * dx, n_per_side + 8, layers)
        nrm = [0.0, 1.0]
    else:
        X = box(dx, -4 * dx, -4 * dx, 0.0 - layers * dx, n_per_side + 8, n_per_side + 8, layers)
        nrm = [0.0, 0.0, 1.0]
    w = WallParticles(kind, dim, X.shape[0])
    w.x_vec_ = X
    w.normal_vec_ = np.tile(nrm, (X.shape[0], 1))
    w.gap_ = np.full(X.shape[0], dx)
    if v is not None:
        w.v_vec_ = np.tile(v, (X.shape[0], 1))
    if t is not None:
        w.t_ = np.full(X.shape[0], t)
    if kappa is not None:
        w.kappa_ = np.full(X.shape[0], kappa)
    return w


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


__all__ = ["lattice_scene", "interior", "random_blob", "wall_plane", "rel_l2", "AABBCheck"]
