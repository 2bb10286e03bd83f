"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle on
identical seeded inputs.

Bars: kernel values and neighbour pair sets bit-exact; fields after N steps
within a relative L2 tolerance (FP64: 1e-9 unless stated -- the only
differences are summation order and ulp-level library functions; FP32_BULK:
the drift bound stated per test)."""
import math
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.particles import CommonLiquidParticle, CompulsiveWallParticle
from ethersph.solve import AABBCheck, Scene, solve
from ethersph.sph_kernel import CubicSpline, SmoothKernel, WendlandC2, kernelGradient, kernelValue
from oracle import oracle as O

from helpers import lattice_scene, random_blob, rel_l2, wall_plane

pytestmark = pytest.mark.gpu

FIELDS = {"x": _lib.FIELD_X, "v": _lib.FIELD_V, "rho": _lib.FIELD_RHO, "p": _lib.FIELD_P, "c": _lib.FIELD_C}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare_fields(dev, ref, tol, thermal=False, fields=None):
    fields = dict(fields or FIELDS)
    if thermal:
        fields["t"] = _lib.FIELD_T
    errs = {}
    assert dev.fluid_count() == ref.fluid_count()
    assert np.array_equal(dev.get(_lib.FIELD_ID), ref.get(_lib.FIELD_ID))
    for name, fid in fields.items():
        errs[name] = rel_l2(dev.get(fid), ref.get(fid))
    say("rel-L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"fields beyond {tol}: {bad}"
    return errs


@pytest.mark.parametrize("kernel", [0, 1, 2, 3])
@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_eval_matches_oracle(ctx, kernel, dim):
    say("kernel_eval", kernel, dim)
    R = 0.037
    r = np.concatenate([np.linspace(0, R * 1.01, 2001), [R, R / 2, R / 3]])
    w = np.empty_like(r)
    dw = np.empty_like(r)
    ctx.check(ctx._L.esph_kernel_eval(ctx.handle, kernel, dim, R, r.size, _lib.dptr(r), _lib.dptr(w),
                                      _lib.dptr(dw)))
    w0, dw0, _ = O.kernel_eval(kernel, dim, R, r)
    if kernel == 1:  # exp(): device and host libm may differ by an ulp
        assert np.allclose(w, w0, rtol=4e-16, atol=0) and np.allclose(dw, dw0, rtol=4e-16, atol=0)
    else:
        assert np.array_equal(w, w0) and np.array_equal(dw, dw0)


def test_kernelValue_api(ctx):
    k = SmoothKernel(1.0, 2, CubicSpline)
    assert math.isclose(kernelValue(0.5, k, ctx), k.sigma_ / 4.0, rel_tol=1e-15)
    assert math.isclose(kernelGradient(0.5, k, ctx), -3.0 * k.sigma_ / 4.0 / k.h_, rel_tol=1e-15)
    assert kernelValue(1.0, k, ctx) == 0.0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("two_sets", [False, True])
def test_neighbor_list_random(ctx, dim, two_sets):
    from ethersph.neighbour import neighborlist

    say("neighbor_list", dim, two_sets)
    rng = np.random.default_rng(10 * dim + two_sets)
    n = 20000 if dim == 2 else 30000
    X = rng.uniform(-1, 1, (n, dim))
    R = 0.02 if dim == 2 else 0.08
    Y = rng.uniform(-1, 1, (n // 2, dim)) if two_sets else None
    a = neighborlist(X, R, Y, ctx=ctx)
    b = O.neighbor_list(X, R, Y, nthreads=8)
    assert a.shape == b.shape and np.array_equal(a, b)


def test_neighbor_list_345(ctx):
    from ethersph.neighbour import neighborlist

    a = neighborlist(np.array([[3.0, 4.0], [0.0, 0.0]]), 10.0, ctx=ctx)
    assert a.tolist() == [[0, 1]]
    assert neighborlist(np.array([[3.0, 4.0], [0.0, 0.0]]), 4.999999, ctx=ctx).shape == (0, 2)


def test_neighbor_list_empty_and_single(ctx):
    from ethersph.neighbour import neighborlist

    assert neighborlist(np.zeros((0, 2)), 1.0, ctx=ctx).shape == (0, 2)
    assert neighborlist(np.zeros((1, 3)), 1.0, ctx=ctx).shape == (0, 2)
    assert neighborlist(np.zeros((1, 2)), 1.0, np.zeros((0, 2)) + 5, ctx=ctx).shape == (0, 2)


def _pairs_equal(dev, ref):
    for ps in (_lib.PAIRS_FLUID_FLUID, _lib.PAIRS_FLUID_WALL):
        a, b = dev.neighbours(ps), ref.neighbours(ps)
        say("pairset", ps, a.shape, b.shape)
        assert a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32_BULK])
def test_scene_pairs_dam_break_2d_full_lattice(ctx, prec):
    """The 5,000-particle dam break: lattice pairs at exactly R = 3 dx."""
    sc = scenarios.collapse_dry_2d()
    say("pairs collapse_dry_2d prec", prec)
    _pairs_equal(sc.make_scene(prec, ctx), O.from_scenario(sc))


@pytest.mark.parametrize("seed", [1, 2])
def test_scene_pairs_dam_break_2d_perturbed(ctx, seed):
    sc = scenarios.perturb(scenarios.collapse_dry_2d(), 0.4, seed)
    say("pairs perturbed seed", seed)
    _pairs_equal(sc.make_scene(_lib.FP64, ctx), O.from_scenario(sc))


def test_scene_pairs_3d(ctx):
    sc = scenarios.perturb(scenarios.dam_break_3d(dx=0.04), 0.3, 5)
    say("pairs 3d", sc.n_fluid, sc.n_wall)
    _pairs_equal(sc.make_scene(_lib.FP64, ctx), O.from_scenario(sc))


def test_fp64_dam_break_2d_parity(ctx):
    """2D dam break at dx = 0.04 (1,250 fluid): 240 steps, Shepard every 30,
    check every 30 (FluidWallSolve)."""
    sc = scenarios.collapse_dry_2d(dx=0.04)
    say("fp64 dam break 2d", sc.n_fluid)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(240)
    dev.sync()
    ref.step(240)
    compare_fields(dev, ref, 1e-9)


def test_fp64_dam_break_2d_full_size(ctx):
    """The 5,000-particle FP64 workload, 2,000 steps (Shepard every 30)."""
    sc = scenarios.collapse_dry_2d()
    say("fp64 dam break 2d full", sc.n_fluid)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(2000)
    dev.sync()
    ref.step(2000, 8)
    compare_fields(dev, ref, 1e-6)


def test_fp64_lid_driven_cavity_parity(ctx):
    """FluidVelocitySolve: static-velocity walls, wall pressure term."""
    sc = scenarios.lid_driven_cavity_2d(dx=0.025, reynolds=100.0)
    say("fp64 lid cavity", sc.n_fluid, sc.n_wall)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(200)
    dev.sync()
    ref.step(200)
    compare_fields(dev, ref, 1e-9)


def test_fp64_thermal_cavity_parity(ctx):
    """FluidCompulsiveThermostaticWallSolve: conduction FF and FW, Boussinesq."""
    sc = scenarios.thermal_cavity_2d(dx=0.025, seed=3)
    say("fp64 thermal cavity", sc.n_fluid, sc.n_wall)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(200)
    dev.sync()
    ref.step(200)
    compare_fields(dev, ref, 1e-9, thermal=True)


def test_fp64_thermostatic_wall_solver_parity(ctx):
    """FluidThermostaticWallSolve: one thermostatic wall group."""
    sc = lattice_scene(n=20, dx=0.01, kernel="WendlandC4", thermal=True, gamma=7, c_0=20.0, g=[0.0, -1.0],
                       reinit=10, dt=2e-5)
    sc.fluid.t_ = np.linspace(0.0, 1.0, sc.n_fluid)
    sc.walls = [wall_plane(2, 20, 0.01, kind=_lib.WALL_THERMOSTATIC, t=1.0, kappa=2.0)]
    sc.model.beta_ = 0.1
    assert sc.solver == _lib.SOLVER_FLUID_THERMOSTATIC_WALL
    say("fp64 thermostatic wall", sc.n_fluid)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(100)
    dev.sync()
    ref.step(100)
    compare_fields(dev, ref, 1e-9, thermal=True)


def test_fp64_dam_break_3d_parity(ctx):
    sc = scenarios.dam_break_3d(dx=0.05)
    say("fp64 dam break 3d", sc.n_fluid, sc.n_wall)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(90)
    dev.sync()
    ref.step(90, 8)
    compare_fields(dev, ref, 1e-9)


@pytest.mark.parametrize("kernel", ["CubicSpline", "Gaussian", "WendlandC2", "WendlandC4"])
def test_fp64_random_blob_all_kernels(ctx, kernel):
    sc = random_blob(n=900, kernel=kernel, reinit=7, walls=[wall_plane(2, 30, 0.01)])
    say("fp64 blob", kernel)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(50)
    dev.sync()
    ref.step(50)
    compare_fields(dev, ref, 1e-9)


def test_fp32_bulk_drift_dam_break_2d(ctx):
    """FP32_BULK against the FP64 oracle after 200 steps of the 5,000-particle
    dam break: every field within 1e-4 relative L2."""
    sc = scenarios.collapse_dry_2d()
    say("fp32 drift 2d")
    dev = sc.make_scene(_lib.FP32_BULK, ctx)
    ref = O.from_scenario(sc)
    dev.step(200)
    dev.sync()
    ref.step(200, 8)
    compare_fields(dev, ref, 1e-4)


@pytest.mark.parametrize("dim,n", [(2, 40), (3, 14)])
def test_fp32_bulk_pairs_near_R_follow_fp64_rule(ctx, dim, n):
    """Gaussian lattice with h = dx (1 + 1e-7) and random velocities above a
    3-layer wall plane: many fluid-fluid and fluid-wall pairs lie 1e-7
    (relative) inside R = 3 h -- inside the band where the FP32_BULK walks'
    quantized float distance cannot decide, so the FP64 rule and binary64
    geometry must.  Dropping those pairs (the oracle with R 2e-7 smaller)
    moves drho by 9e-4 (2D) / 2e-3 (3D) and dv by 1.1e-3 / 2.4e-3 relative
    L2; after one step the device's drho and dv (continuity pass, and the
    force pass's fluid and wall walks) stay within 5e-5 of the FP64 oracle's
    and the positions' displacement within 1e-6.  The walks' own decisions
    equal the oracle's pair sets.  (Exactly at R the truncated Gaussian jumps
    to zero, and a float q = r/h that rounds to 3 drops the pair from the sum
    whatever the rule says; the lattice is therefore kept just inside R.)"""
    sc = lattice_scene(n=n, dx=0.01, dim=dim, kernel="Gaussian", h_factor=1.0 + 1e-7)
    rng = np.random.default_rng(2)
    sc.fluid.v_vec_ = rng.uniform(-0.5, 0.5, sc.fluid.x_vec_.shape)
    sc.walls = [wall_plane(dim, n, 0.01)]
    r0 = sc.fluid.rho_.copy()
    v0 = sc.fluid.v_vec_.copy()
    x0 = sc.fluid.x_vec_.copy()
    dev = sc.make_scene(_lib.FP32_BULK, ctx)
    ref = O.from_scenario(sc)
    ff = ref.neighbours(_lib.PAIRS_FLUID_FLUID)
    fw = ref.neighbours(_lib.PAIRS_FLUID_WALL)
    w = dev.walk_pairs(_lib.PAIRS_FLUID_FLUID)
    assert np.array_equal(w[w[:, 0] < w[:, 1]], ff) and len(w) == 2 * len(ff)
    assert np.array_equal(dev.walk_pairs(_lib.PAIRS_FLUID_WALL), fw)
    dev.step(1)
    dev.sync()
    ref.step(1)
    errs = {"drho": rel_l2(dev.get(_lib.FIELD_RHO) - r0, ref.get(_lib.FIELD_RHO) - r0),
            "dv": rel_l2(dev.get(_lib.FIELD_V) - v0, ref.get(_lib.FIELD_V) - v0),
            "dx": rel_l2(dev.get(_lib.FIELD_X) - x0, ref.get(_lib.FIELD_X) - x0)}
    say("fp32 pairs near R", dim, errs)
    assert errs["drho"] < 5e-5 and errs["dv"] < 5e-5 and errs["dx"] < 1e-6, errs


def test_check_deletion_matches_oracle(ctx):
    """Particles thrown out of the AABB are deleted at check steps, stably."""
    sc = lattice_scene(n=16, dx=0.01, check_interval=5, check=AABBCheck(lo=[-1.0, 0.02], hi=[0.12, 1.0]),
                       g=[3000.0, -2000.0], dt=1e-4)
    say("check deletion")
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    for _ in range(4):
        dev.step(5)
        ref.step(5)
        assert dev.fluid_count() == ref.fluid_count()
    assert ref.fluid_count() < sc.n_fluid
    compare_fields(dev, ref, 1e-9)
    _pairs_equal(dev, ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_check_planes_match_oracle(ctx, dim):
    """Box plus tilted half-spaces: the same particles are deleted, stably."""
    planes = [((0.6, 0.8, 0.0)[:dim], 0.11), ((-0.5, 0.3, 0.8)[:dim], 0.02)]
    g = [3000.0, -2000.0] if dim == 2 else [3000.0, -2000.0, 1500.0]
    sc = lattice_scene(n=16 if dim == 2 else 8, dx=0.01, dim=dim, check_interval=5,
                       check=AABBCheck(lo=[-1.0] * dim, hi=[0.2] * dim, planes=planes), g=g, dt=1e-4)
    say("check planes", dim)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    for _ in range(4):
        dev.step(5)
        ref.step(5)
        assert dev.fluid_count() == ref.fluid_count()
    assert ref.fluid_count() < sc.n_fluid
    compare_fields(dev, ref, 1e-9)
    _pairs_equal(dev, ref)


def test_nan_position_is_reported(ctx):
    sc = lattice_scene(n=8, dx=0.01)
    sc.fluid.x_vec_[5, 0] = np.nan
    say("nan position")
    dev = sc.make_scene(_lib.FP64, ctx)
    dev.step(2)
    with pytest.raises(_lib.EsphError) as e:
        dev.sync()
    assert e.value.code == 4
    assert dev.fluid_count() == sc.n_fluid  # the reference keeps such a particle


def test_reupload_capacity_and_errors(ctx):
    sc = lattice_scene(n=10, dx=0.01)
    dev = sc.make_scene(_lib.FP64, ctx, capacity=150)
    dev.step(3)
    dev.set_fluid(sc.fluid)
    assert dev.step_count == 0 and dev.fluid_count() == 100
    big = lattice_scene(n=13, dx=0.01)
    with pytest.raises(_lib.EsphError) as e:
        dev.set_fluid(big.fluid)
    assert e.value.code == 6
    bad = sc.desc()
    bad.dim = 4
    with pytest.raises(_lib.EsphError):
        Scene(bad, sc.fluid, [], ctx)


def test_solve_dropin_particle_lists(ctx):
    """solve(fluid, walls, kernel, model, td; check) on particle objects
    mutates them like solve! does (including deleteat!)."""
    sc = lattice_scene(n=12, dx=0.01, check_interval=10, check=AABBCheck(hi=[0.1, 1.0]), g=[200.0, -9.81],
                       dt=1e-4, reinit=10)
    sc.td.total_step_ = 40
    fluid = []
    for i in range(sc.n_fluid):
        p = CommonLiquidParticle(2)
        p.x_vec_[:] = sc.fluid.x_vec_[i]
        p.rho_, p.mass_, p.gap_, p.c_ = 1000.0, sc.fluid.mass_[i], 0.01, 10.0
        fluid.append(p)
    w = wall_plane(2, 12, 0.01)
    walls = []
    for i in range(len(w)):
        q = CompulsiveWallParticle(2)
        q.x_vec_[:] = w.x_vec_[i]
        q.normal_vec_[:] = w.normal_vec_[i]
        q.gap_ = 0.01
        walls.append(q)
    sc.walls = [w]
    ref = O.from_scenario(sc)
    ref.step(40)
    solve(fluid, walls, sc.kernel, sc.model, sc.td, check=AABBCheck(hi=[0.1, 1.0]), ctx=ctx)
    assert len(fluid) == ref.fluid_count() < 144
    X = np.array([p.x_vec_ for p in fluid])
    assert rel_l2(X, ref.get(_lib.FIELD_X)) < 1e-9


def test_overflow_scan_path_matches_oracle(ctx):
    """Particles packed at 0.45 of the lattice spacing overflow the neighbour
    lists (sized from the gap); they take the cell-scan path."""
    sc = random_blob(n=900, kernel="WendlandC2", reinit=5, spacing=0.01)
    sc.fluid.x_vec_ = sc.fluid.x_vec_ * 0.45
    say("overflow path")
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(20)
    dev.sync()
    ref.step(20)
    assert dev.stats()["list_overflow"] > 0
    compare_fields(dev, ref, 1e-9)


@pytest.mark.parametrize("skin", [-1.0, 0.0, 0.3])
def test_candidate_list_reuse_keeps_parity(ctx, skin):
    """Candidate lists reused across steps (skin 0.045 by default in 2D, 0.3) or
    rebuilt every step (negative skin): the same pair sets and fields as the
    oracle; with reuse, fewer sorts than steps."""
    sc = scenarios.collapse_dry_2d(dx=0.04)
    d = sc.desc(_lib.FP64)
    d.verlet_skin = skin
    dev = Scene(d, sc.fluid, sc.walls, ctx)
    ref = O.from_scenario(sc)
    for chunk in (37, 63):
        dev.step(chunk)
        ref.step(chunk)
        dev.sync()
        _pairs_equal(dev, ref)
    compare_fields(dev, ref, 1e-9)
    st = dev.stats()
    say("skin", skin, st)
    if skin < 0:
        assert st["sorts"] == 101 and not st["reused_lists"]
    else:
        assert st["reused_lists"] and st["sorts"] < 101
    dev.close()
