"""GPU parity of the inlet/outlet solver (FluidInletOutletWallSolve.jl)
against the FP64 oracle: migrations (counts, ids, order) exact; isolated
particles bit-exact; Poiseuille channel and pipe fields within 1e-9 relative
L2 (FP64) and 1e-4 (FP32_BULK); buffer scenes inside a batch and the server
equal to the same scene stepped alone."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.serving import Server
from ethersph.solve import Batch, solve
from oracle import oracle as O

from helpers import free_io_scene, rel_l2

pytestmark = pytest.mark.gpu

GROUPS = (_lib.GROUP_FLUID, _lib.GROUP_INLET, _lib.GROUP_OUTLET)
FIELDS = {"x": _lib.FIELD_X, "v": _lib.FIELD_V, "rho": _lib.FIELD_RHO, "p": _lib.FIELD_P, "c": _lib.FIELD_C}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare_groups(dev, ref, tol):
    errs = {}
    for g in GROUPS:
        n = dev.group_count(g)
        assert n == ref.group_count(g), ("count", g)
        if n == 0:
            continue
        assert np.array_equal(dev.get(_lib.FIELD_ID, g), ref.get(_lib.FIELD_ID, g)), ("ids", g)
        for name, fid in FIELDS.items():
            a, b = dev.get(fid, g), ref.get(fid, g)
            errs[(g, name)] = 0.0 if tol == 0 and np.array_equal(a, b) else rel_l2(a, b)
    say("rel-L2:", {f"{g}.{k}": f"{v:.2e}" for (g, k), v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"fields beyond {tol}: {bad}"


@pytest.mark.parametrize("steps", [1, 11, 21, 50, 60, 160])
def test_free_particles_migrate_bit_exact(ctx, steps):
    say("free io", steps)
    sc = free_io_scene()
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(steps)
    dev.sync()
    ref.step(steps)
    compare_groups(dev, ref, 0.0)
    dev.close()


def test_poiseuille_2d_fp64_parity(ctx):
    sc = scenarios.poiseuille_2d(rows=12, length_factor=2)
    say("poiseuille 2d", sc.n_fluid, [len(b) for b in sc.buffers], sc.n_wall)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    for pairset in (_lib.PAIRS_FLUID_FLUID, _lib.PAIRS_FLUID_WALL):
        assert np.array_equal(dev.neighbours(pairset), ref.neighbours(pairset))
    for chunk in (150, 250, 400):
        dev.step(chunk)
        ref.step(chunk, 4)
        dev.sync()
        compare_groups(dev, ref, 1e-9)
        for pairset in (_lib.PAIRS_FLUID_FLUID, _lib.PAIRS_FLUID_WALL):
            assert np.array_equal(dev.neighbours(pairset), ref.neighbours(pairset))
    assert ref.group_count(_lib.GROUP_FLUID) != sc.n_fluid or ref.get(_lib.FIELD_ID).max() >= sc.n_fluid
    dev.close()


def test_poiseuille_3d_fp64_parity(ctx):
    sc = scenarios.poiseuille_3d(cells=10, length_factor=1)
    say("poiseuille 3d", sc.n_fluid, [len(b) for b in sc.buffers], sc.n_wall)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    dev.step(120)
    ref.step(120, 4)
    dev.sync()
    compare_groups(dev, ref, 1e-9)
    assert np.array_equal(dev.neighbours(_lib.PAIRS_FLUID_FLUID), ref.neighbours(_lib.PAIRS_FLUID_FLUID))
    dev.close()


def test_poiseuille_2d_fp32_bulk_drift(ctx):
    """FP32 pair arithmetic, FP64 integration: 400 steps within 1e-4."""
    sc = scenarios.poiseuille_2d(rows=12, length_factor=2)
    dev = sc.make_scene(_lib.FP32_BULK, ctx)
    ref = O.from_scenario(sc)
    dev.step(400)
    ref.step(400, 4)
    dev.sync()
    compare_groups(dev, ref, 1e-4)
    dev.close()


def test_buffer_scenes_in_batch_equal_alone(ctx):
    scs = [scenarios.poiseuille_2d(rows=10, length_factor=2), scenarios.collapse_dry_2d(dx=0.05),
           scenarios.poiseuille_3d(cells=8, length_factor=1), free_io_scene()]
    for prec in (_lib.FP64, _lib.FP32_BULK):
        say("io batch", prec)
        alone = [s.make_scene(prec, ctx) for s in scs]
        mixed = [s.make_scene(prec, ctx) for s in scs]
        for s in alone:
            s.step(90)
        b = Batch(mixed, graph=True)
        b.step(50)
        b.step(40)
        b.sync()
        for a, m in zip(alone, mixed):
            a.sync()
            for g in GROUPS:
                assert a.group_count(g) == m.group_count(g)
                if a.group_count(g):
                    for fid in list(FIELDS.values()) + [_lib.FIELD_ID]:
                        assert np.array_equal(a.get(fid, g), m.get(fid, g)), (g, fid)
        b.close()
        for s in alone + mixed:
            s.close()


def test_server_runs_buffer_requests(ctx):
    sv = Server(ctx)
    p2 = scenarios.poiseuille_2d(rows=10, length_factor=2)
    p3 = scenarios.poiseuille_3d(cells=8, length_factor=1)
    a = sv.submit_scenario(p2, _lib.FP32_BULK, 70)
    sv.advance(20)
    b = sv.submit_scenario(p3, _lib.FP32_BULK, 45)
    for _ in range(10):
        if all(sv.status(r)[0] in (_lib.REQ_DONE, _lib.REQ_FAILED) for r in (a, b)):
            break
        sv.advance(30)
    assert [sv.status(r)[0] for r in (a, b)] == [_lib.REQ_DONE, _lib.REQ_DONE]
    for rid, sc, n in ((a, p2, 70), (b, p3, 45)):
        s = sc.make_scene(_lib.FP32_BULK, ctx)
        s.step(n)
        s.sync()
        for fid in (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_ID):
            assert np.array_equal(sv.result(rid, fid), s.get(fid)), (sc.name, fid)
        s.close()
        sv.release(rid)
    sv.close()


def test_solve_inlet_outlet_form(ctx):
    """solve!(fluid, inlet, outlet, walls, kernel, model, td; closures as
    data): the caller's groups are replaced by the state after total_step_."""
    sc = scenarios.poiseuille_2d(rows=10, length_factor=2, total_step=130)
    ref = O.from_scenario(sc)
    ref.step(130)
    fluid = sc.fluid.subset(np.arange(len(sc.fluid)))
    inlet, outlet = (b.subset(np.arange(len(b))) for b in sc.buffers)
    solve(fluid, inlet, outlet, sc.walls[0], sc.kernel, sc.model, sc.td, inlet_outlet=sc.planes,
          user_defined_function=sc.profile, ctx=ctx)
    for grp, g in ((fluid, _lib.GROUP_FLUID), (inlet, _lib.GROUP_INLET), (outlet, _lib.GROUP_OUTLET)):
        assert len(grp) == ref.group_count(g)
        assert rel_l2(grp.x_vec_, ref.get(_lib.FIELD_X, g)) < 1e-9
        assert rel_l2(grp.v_vec_, ref.get(_lib.FIELD_V, g)) < 1e-9


def test_capacity_exhaustion_is_reported(ctx):
    """A capacity with no room for inlet copies: the copies are dropped and
    sync reports ESPH_ERR_CAPACITY."""
    sc = free_io_scene()
    s = sc.make_scene(_lib.FP64, ctx, capacity=sc.n_particles)
    s.step(20)
    s.sync()
    s.step(5)  # two inlet particles enter at step 21
    with pytest.raises(_lib.EsphError) as e:
        s.sync()
    assert e.value.code == 6  # ESPH_ERR_CAPACITY
    s.close()
