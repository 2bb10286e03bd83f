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


def free_io_scene(u=1.0, dt=1e-3, R=0.03, profile=None):
    """Inlet/outlet scene of isolated particles (pairwise farther apart than
    R at every step, no body force): every particle moves at its constant
    velocity, so the migrations follow from the positions alone.

    inlet_enter = 0, inlet_shift = 0.1, outlet_enter = 1.0, outlet_exit = 1.02
    along x; u * dt = 1e-3 per step."""
    from ethersph.solve import InletOutletPlanes

    dim = 2
    sk = SmoothKernel(R, dim, KERNELS["CubicSpline"])
    model = CommonWeaklyCompressibleLiquidModel(1000.0, 10.0, 0.0, 1, 1e-3, [0.0, 0.0])
    td = DensityReinitializedFixedTimeIntervalForwardEuler(dt, 1, 1, 0, 1000)

    def group(X, V):
        p = LiquidParticles(dim, len(X))
        p.x_vec_ = np.asarray(X, dtype=np.float64)
        p.v_vec_ = np.asarray(V, dtype=np.float64)
        p.rho_ = np.full(len(p), 1000.0)
        p.c_ = np.full(len(p), 10.0)
        p.mass_ = np.full(len(p), 0.1)
        p.gap_ = np.full(len(p), 0.01)
        return p

    # fluid: two leave together at step 50 (x0 + 50 u dt > 1.0), one at 60
    fluid = group([[0.9505, 0.50], [0.9405, 0.60], [0.9505, 0.40], [0.3, 0.9]],
                  [[u, 0.0], [u, 0.0], [u, 0.0], [0.0, 0.0]])
    # inlet: two enter together at step 21, one at 31; each re-enters after
    # 100 more steps (shift 0.1)
    inlet = group([[-0.0205, 0.20], [-0.0305, 0.10], [-0.0205, 0.05]], [[u, 0.0], [u, 0.0], [u, 0.0]])
    # outlet: one leaves at step 11 (x > 1.02), one stays put
    outlet = group([[1.0105, 0.70], [1.01, 0.85]], [[u, 0.0], [0.0, 0.0]])
    w = WallParticles(_lib.WALL_COMPULSIVE, dim, 1)
    w.x_vec_ = np.array([[5.0, 5.0]])
    w.normal_vec_ = np.array([[0.0, 1.0]])
    w.gap_ = np.array([0.01])
    planes = InletOutletPlanes(0, 0.0, 0.1, 1.0, 1.02)
    return Scenario("free_io", dim, sk, model, td, fluid, [w], None, (inlet, outlet), planes, profile)


def migrate_by_lists(fluid, inlet, outlet, steps, dt, planes):
    """The migration rules of FluidInletOutletWallSolve.jl:193-262 on plain
    Python lists of free particles (x, v, id): per step every particle moves
    x += v dt; then the three predicates are evaluated on the moved
    positions; inlet copies are push!ed onto the fluid in inlet order, the
    leaving fluid particles push!ed onto the outlet in fluid order, the
    leavers and the expired outlet particles removed with deleteat!, and the
    entering inlet particles moved back."""
    fl = [[np.array(x, dtype=np.float64), np.array(v, dtype=np.float64), i]
          for i, (x, v) in enumerate(zip(fluid.x_vec_, fluid.v_vec_))]
    il = [[np.array(x, dtype=np.float64), np.array(v, dtype=np.float64), i]
          for i, (x, v) in enumerate(zip(inlet.x_vec_, inlet.v_vec_))]
    ol = [[np.array(x, dtype=np.float64), np.array(v, dtype=np.float64), i]
          for i, (x, v) in enumerate(zip(outlet.x_vec_, outlet.v_vec_))]
    next_f, next_o = len(fl), len(ol)
    a = planes.axis
    for _ in range(steps):
        for p in fl + il + ol:
            p[0] = p[0] + p[1] * dt - 0.0 * (dt * dt) / 2.0
        enter = [k for k, p in enumerate(il) if p[0][a] > planes.inlet_enter]
        leave = [k for k, p in enumerate(fl) if p[0][a] > planes.outlet_enter]
        gone = [k for k, p in enumerate(ol) if p[0][a] > planes.outlet_exit]
        for k in enter:
            fl.append([il[k][0].copy(), il[k][1].copy(), next_f])
            next_f += 1
        for k in leave:
            ol.append([fl[k][0].copy(), fl[k][1].copy(), next_o])
            next_o += 1
        fl = [p for k, p in enumerate(fl) if k not in set(leave)]
        ol = [p for k, p in enumerate(ol) if k not in set(gone)]
        for k in enter:
            il[k][0][a] -= planes.inlet_shift
    return fl, il, ol
