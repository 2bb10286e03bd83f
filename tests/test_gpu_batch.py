"""GPU: batched (ensemble) stepping and graph replay give the same result as
stepping each scene alone.  Each particle's arithmetic does not depend on
which other scenes share the launch, so the bar is bit-identical fields."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import Batch

from helpers import lattice_scene, wall_plane

pytestmark = pytest.mark.gpu

FIELDS = (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_C)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_scenarios():
    return [
        scenarios.collapse_dry_2d(dx=0.04),
        scenarios.lid_driven_cavity_2d(dx=0.03, reynolds=400.0),
        scenarios.thermal_cavity_2d(dx=0.03, seed=7),
        scenarios.dam_break_3d(dx=0.06),
        scenarios.collapse_dry_2d(dx=0.05, kernel=scenarios.WendlandC2),
        scenarios.lid_driven_cavity_2d(dx=0.04, kernel=scenarios.CubicSpline, reynolds=50.0),
    ]


def assert_same(a, b):
    assert a.fluid_count() == b.fluid_count()
    for f in FIELDS + ((_lib.FIELD_T,) if a.thermal else ()):
        assert np.array_equal(a.get(f), b.get(f)), f"field {f} differs"


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32_BULK])
def test_batch_equals_alone(ctx, prec):
    scs = mixed_scenarios()
    say("batch vs alone, prec", prec, [s.n_fluid for s in scs])
    alone = [s.make_scene(prec, ctx) for s in scs]
    mixed = [s.make_scene(prec, ctx) for s in scs]
    for s in alone:
        s.step(65)
    b = Batch(mixed, graph=True)
    b.step(40)
    b.step(25)
    b.sync()
    for a, m in zip(alone, mixed):
        a.sync()
        assert m.step_count == 65
        assert_same(a, m)
    b.close()


def test_batch_without_graph_and_members_stepped_alone(ctx):
    scs = mixed_scenarios()[:3]
    alone = [s.make_scene(_lib.FP64, ctx) for s in scs]
    mixed = [s.make_scene(_lib.FP64, ctx) for s in scs]
    for s in alone:
        s.step(50)
    b = Batch(mixed, graph=False)
    b.step(20)
    mixed[1].step(10)  # a member can still be stepped alone
    b.step(30)
    mixed[0].step(10)
    mixed[2].step(10)
    alone[1].step(10)
    alone[0].step(10)
    alone[2].step(10)
    b.sync()
    for a, m in zip(alone, mixed):
        assert m.step_count == a.step_count == 60
        assert_same(a, m)
    b.close()


def test_scene_graph_equals_plain(ctx):
    """95 steps replayed from step graphs (single steps with and without the
    Shepard pass, and runs of 2-16 steps in one graph) against plain launches."""
    sc = scenarios.collapse_dry_2d(dx=0.03)
    a = sc.make_scene(_lib.FP32_BULK, ctx)
    g = sc.make_scene(_lib.FP32_BULK, ctx)
    g.enable_graph(True)
    a.step(95)
    g.step(60)
    g.step(35)
    a.sync()
    g.sync()
    assert g.step_count == 95
    assert_same(a, g)


def test_batch_with_deletion_and_overflow_paths(ctx):
    """A scene whose particles leave the check box (deletion) and a crowded
    scene whose neighbour lists overflow (cell-scan path) inside a batch."""
    from ethersph.solve import AABBCheck

    s1 = lattice_scene(n=16, dx=0.01, check_interval=5, check=AABBCheck(lo=[-1.0, 0.02], hi=[0.12, 1.0]),
                       g=[3000.0, -2000.0], dt=1e-4)
    s2 = lattice_scene(n=20, dx=0.01, kernel="CubicSpline")
    # crowd the second scene: particles at a quarter of the lattice spacing
    s2.fluid.x_vec_ = s2.fluid.x_vec_ * 0.25
    s2.fluid.gap_[:] = 0.01
    scs = [s1, s2]
    alone = [s.make_scene(_lib.FP64, ctx) for s in scs]
    mixed = [s.make_scene(_lib.FP64, ctx) for s in scs]
    for s in alone:
        s.step(20)
    b = Batch(mixed, graph=True)
    b.step(20)
    b.sync()
    assert alone[1].stats()["list_overflow"] > 0
    for a, m in zip(alone, mixed):
        assert_same(a, m)


def test_graph_shepard_schedule_with_members_out_of_phase(ctx):
    """Step graphs launch the Shepard pass only on steps where some member
    re-initialises its density (chosen on the host from the members' step
    counts).  Members with different intervals (30 and 5, the second on the
    inlet/outlet solver) and step counts out of phase, replayed as one
    batch, match each scene stepped alone."""
    scs = [scenarios.lid_driven_cavity_2d(dx=0.04, reynolds=100.0),
           scenarios.poiseuille_2d(rows=8, reynolds=10.0)]
    say("shepard schedule", [s.n_fluid for s in scs])
    alone = [s.make_scene(_lib.FP32_BULK, ctx) for s in scs]
    mixed = [s.make_scene(_lib.FP32_BULK, ctx) for s in scs]
    mixed[0].step(7)
    mixed[1].step(3)
    b = Batch(mixed, graph=True)
    b.step(33)
    b.step(4)
    b.sync()
    alone[0].step(44)
    alone[1].step(40)
    for a, m in zip(alone, mixed):
        a.sync()
        assert m.step_count == a.step_count
        assert_same(a, m)
    b.close()
