"""GPU: the force pass's shared wall walk (the default: the wall candidates
of a wavefront's 64 particles shared out over all its lanes, each particle's
wall terms summed by a segmented sum across the lanes) against each particle
walking its own wall candidates (Context.set_wall_lanes).  Both add the same
terms, in a different order, so the fields agree to rounding: after 20 steps
within 1e-10 relative L2 in FP64 and 1e-4 in FP32_BULK.  A mistake in the
sharing (a candidate taken twice or dropped, a term credited to the wrong
particle, a segment cut wrongly) moves the near-wall particles by orders of
magnitude more.  Covered: 2D and 3D, both precisions, moving walls (lid),
a cavity whose walls touch every wavefront, buffer particles of the
inlet/outlet solver, overflowed lists (cell-scan path beside the shared
walk), and a batch replayed as a graph."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import Batch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = (("x", _lib.FIELD_X), ("v", _lib.FIELD_V), ("rho", _lib.FIELD_RHO), ("p", _lib.FIELD_P))
BAR = {_lib.FP64: 1e-10, _lib.FP32_BULK: 1e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctxs():
    shared = _lib.Context(0)
    lanes = _lib.Context(0)
    lanes.set_wall_lanes(True)
    yield shared, lanes
    shared.close()
    lanes.close()


def crowded_corner():
    """2D dry collapse with the water packed to 0.45 of its spacing into the
    corner: its lists (sized from the gap) overflow next to the walls."""
    sc = scenarios.collapse_dry_2d(dx=0.04)
    sc.fluid.x_vec_ = sc.fluid.x_vec_ * 0.45
    return sc


CASES = {
    "dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.03),
    "collapse_dry_2d": lambda: scenarios.collapse_dry_2d(dx=0.02),
    "lid_driven_cavity_2d": lambda: scenarios.lid_driven_cavity_2d(dx=0.02, reynolds=100.0),
    "poiseuille_2d": lambda: scenarios.poiseuille_2d(rows=8, reynolds=10.0),
    "crowded_corner": crowded_corner,
}


def compare(a, b, prec, what):
    assert a.fluid_count() == b.fluid_count(), what
    assert np.array_equal(a.get(_lib.FIELD_ID), b.get(_lib.FIELD_ID)), what
    errs = {n: rel_l2(a.get(f), b.get(f)) for n, f in FIELDS}
    say(" ", what, {k: f"{v:.1e}" for k, v in errs.items()})
    assert all(v < BAR[prec] for v in errs.values()), (what, errs)


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32_BULK])
@pytest.mark.parametrize("case", sorted(CASES))
def test_shared_wall_walk_matches_per_particle_walk(ctxs, case, prec):
    sc = CASES[case]()
    say("wall walk shared vs per particle:", case, "fluid", sc.n_fluid, "wall", sc.n_wall, "prec", prec)
    shared, lanes = (sc.make_scene(prec, c) for c in ctxs)
    assert len(shared.neighbours(_lib.PAIRS_FLUID_WALL)) > 0
    for s in (shared, lanes):
        s.step(20)
        s.sync()
    if case == "crowded_corner":
        assert shared.stats()["list_overflow"] > 0 and lanes.stats()["list_overflow"] > 0
    compare(shared, lanes, prec, case)
    shared.close()
    lanes.close()


def test_shared_wall_walk_in_a_graph_batch(ctxs):
    """Members of one batch, replayed as a graph, against the same scenes
    stepped alone with per-particle wall walks."""
    shared_ctx, lanes_ctx = ctxs
    scs = [scenarios.dam_break_3d(dx=0.04), scenarios.collapse_dry_2d(dx=0.04),
           scenarios.lid_driven_cavity_2d(dx=0.04, reynolds=100.0)]
    say("graph batch, shared wall walk", [s.n_fluid for s in scs])
    ref = [s.make_scene(_lib.FP32_BULK, lanes_ctx) for s in scs]
    mix = [s.make_scene(_lib.FP32_BULK, shared_ctx) for s in scs]
    for r in ref:
        r.step(30)
        r.sync()
    b = Batch(mix, graph=True)
    b.step(15)
    b.step(15)
    b.sync()
    for k, (r, m) in enumerate(zip(ref, mix)):
        compare(m, r, _lib.FP32_BULK, f"member {k}")
    b.close()
    for s in ref + mix:
        s.close()
