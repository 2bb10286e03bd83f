"""GPU: the force pass's shared wall walk (the default in 3D: the wall
candidates of a wavefront's 64 particles shared out over all its lanes, each
particle's wall terms summed by a segmented sum across the lanes) against
each particle walking its own wall candidates (Context.set_wall_lanes; 2D
scenes always do).  Both add the same terms, in a different order, so the
fields agree to rounding: after 20 steps (6 for a scene that blows up)
within 1e-10 relative L2 in FP64 and 1e-4 in FP32_BULK.  A mistake in the
sharing (a candidate taken twice
or dropped, a term credited to the wrong particle, a segment cut wrongly)
moves the near-wall particles by orders of magnitude more.  Covered (3D,
both precisions): the dam break, a closed box (walls on all six faces), the
Poiseuille pipe with the buffer particles of the inlet/outlet solver,
overflowed lists (cell-scan path beside the shared walk), and a batch
replayed as a graph; one 2D scene checks that 2D is unchanged."""
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


def crowded_block():
    """3D dam break with a block of water, away from the walls and a few
    spacings from the end wall, packed to 0.4 of its spacing: its particles'
    lists (sized from the gap) overflow and take the cell scan, in
    wavefronts whose particles next to the end wall share out their wall
    candidates."""
    sc = scenarios.dam_break_3d(dx=0.04)
    X = sc.fluid.x_vec_
    idx = np.rint(X / 0.04 - 0.5).astype(int)
    sel = np.all((idx >= [4, 8, 4]) & (idx <= [10, 14, 10]), axis=1)
    c = X[sel].mean(axis=0)
    X[sel] = c + (X[sel] - c) * 0.4
    sc.fluid.x_vec_ = X
    return sc


CASES = {
    "dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.03),
    "closed_box_3d": lambda: scenarios.dam_break_3d(dx=0.04, lid=True),
    "poiseuille_3d": lambda: scenarios.poiseuille_3d(cells=8, length_factor=1),
    "crowded_block_3d": crowded_block,
    "collapse_dry_2d": lambda: scenarios.collapse_dry_2d(dx=0.02),
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
    # (the crowded block blows up within a dozen steps -- velocities of 1e5
    # m/s -- and then amplifies any rounding difference: it runs 6)
    steps = 6 if case == "crowded_block_3d" else 20
    for s in (shared, lanes):
        s.step(steps)
        s.sync()
    if case == "crowded_block_3d":
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
