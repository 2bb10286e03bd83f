"""GPU: the pair-pass tuner (gather depth and workgroup size).  Results are
bit-identical for every choice (the pair passes visit the same neighbours in
the same order), the tuned choice is recorded per scene class, and batches
and graphs use it."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import Batch

pytestmark = pytest.mark.gpu

FIELDS = (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_C)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(sc, prec, steps, tune, depth, graph=False, threads=256):
    ctx = _lib.Context(0)
    ctx.set_tuning(tune, depth)
    ctx.set_workgroup(threads)
    s = sc.make_scene(prec, ctx)
    if graph:
        s.enable_graph(True)
    s.step(steps)
    s.sync()
    out = {f: s.get(f) for f in FIELDS + ((_lib.FIELD_T,) if s.thermal else ())}
    mix = 2 if s.thermal else 1
    tuned = ctx.tuned_depth(sc.dim, sc.kernel.kind, prec, mix)
    if tuned is not None:
        tuned = (tuned[0], tuned[1], [ctx.tuned_pass(sc.dim, sc.kernel.kind, prec, q, mix) for q in (0, 1, 2)])
    s.close()
    ctx.close()
    return out, tuned


def dam_break_every_step_shepard():
    """Re-initialisation at every step, so the tuning step (the first) is one:
    the tuner's dry continuity and forced Shepard runs write SceneDev::rhon
    before the step's own Shepard pass and must leave no trace."""
    sc = scenarios.dam_break_3d(dx=0.06)
    sc.td.density_reinitialized_interval_ = 1
    return sc


@pytest.mark.parametrize("name,prec", [("dam_break_3d", _lib.FP32_BULK), ("thermal_cavity_2d", _lib.FP64),
                                       ("poiseuille_2d", _lib.FP32_BULK), ("lid_driven_cavity_2d", _lib.FP64),
                                       ("dam_break_3d_shepard_each_step", _lib.FP32_BULK)])
def test_results_do_not_depend_on_depth(name, prec):
    sc = {"dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.06),
          "thermal_cavity_2d": lambda: scenarios.thermal_cavity_2d(dx=0.03, seed=3),
          "poiseuille_2d": lambda: scenarios.poiseuille_2d(rows=10, length_factor=2),
          "lid_driven_cavity_2d": lambda: scenarios.lid_driven_cavity_2d(dx=0.04),
          "dam_break_3d_shepard_each_step": dam_break_every_step_shepard}[name]()
    say("depth independence", name, prec)
    ref, _ = run(sc, prec, 40, False, 1)
    for depth in (2, 4):
        got, _ = run(sc, prec, 40, False, depth)
        for f in ref:
            assert np.array_equal(got[f], ref[f]), (depth, f)
    for threads in (64, 128):
        got, _ = run(sc, prec, 40, False, 4, threads=threads)
        for f in ref:
            assert np.array_equal(got[f], ref[f]), (threads, f)
    got, tuned = run(sc, prec, 40, True, 2)
    assert tuned is not None and tuned[0] in (1, 2, 4) and all(t > 0 for t in tuned[1])
    for q in (0, 1, 2):  # each pass's own choice, from its own timings
        d, t, ms = tuned[2][q]
        assert d in (1, 2, 4) and t in (64, 128, 256)
        assert all(v > 0 for v in (ms[:6] if q == 0 else ms)), (q, ms)
    say("tuned", tuned)
    for f in ref:
        assert np.array_equal(got[f], ref[f]), ("tuned", f)
    got, _ = run(sc, prec, 40, True, 2, graph=True)
    for f in ref:
        assert np.array_equal(got[f], ref[f]), ("tuned graph", f)


def test_batch_tunes_each_class_once():
    ctx = _lib.Context(0)
    scs = [scenarios.collapse_dry_2d(dx=0.05), scenarios.lid_driven_cavity_2d(dx=0.04),
           scenarios.dam_break_3d(dx=0.08)]
    scenes = [s.make_scene(_lib.FP32_BULK, ctx) for s in scs]
    b = Batch(scenes, graph=True)
    b.step(10)
    b.sync()
    for sc in scs:
        assert ctx.tuned_depth(sc.dim, sc.kernel.kind, _lib.FP32_BULK, 1) is not None
        wg = ctx.tuned_workgroup(sc.dim, sc.kernel.kind, _lib.FP32_BULK, 1)
        assert wg is not None and wg[0] in (64, 128, 256) and all(t > 0 for t in wg[1])
    b.close()
    for s in scenes:
        s.close()
    ctx.close()


def test_graphs_recaptured_when_the_launch_settings_change():
    """A scene in graph mode whose context changes its workgroup size and
    gather depth between calls (its plan then drops every step graph, single
    steps and runs of 2-16 steps, and captures them again with the new
    settings): the fields equal plain launches with fixed settings bit for
    bit."""
    sc = scenarios.collapse_dry_2d(dx=0.03)
    ref, _ = run(sc, _lib.FP32_BULK, 70, False, 2)
    ctx = _lib.Context(0)
    ctx.set_tuning(False, 4)
    ctx.set_workgroup(64)
    s = sc.make_scene(_lib.FP32_BULK, ctx)
    s.enable_graph(True)
    s.step(35)
    ctx.set_tuning(False, 1)
    ctx.set_workgroup(128)
    s.step(35)
    s.sync()
    say("graphs recaptured after a settings change, steps", s.step_count)
    assert s.step_count == 70
    for f in FIELDS:
        assert np.array_equal(s.get(f), ref[f]), ("recaptured graph", f)
    s.close()
    ctx.close()
