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
    tuned = ctx.tuned_depth(sc.dim, sc.kernel.kind, prec, 2 if s.thermal else 1)
    s.close()
    ctx.close()
    return out, tuned


@pytest.mark.parametrize("name,prec", [("dam_break_3d", _lib.FP32_BULK), ("thermal_cavity_2d", _lib.FP64),
                                       ("poiseuille_2d", _lib.FP32_BULK), ("lid_driven_cavity_2d", _lib.FP64)])
def test_results_do_not_depend_on_depth(name, prec):
    sc = {"dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.06),
          "thermal_cavity_2d": lambda: scenarios.thermal_cavity_2d(dx=0.03, seed=3),
          "poiseuille_2d": lambda: scenarios.poiseuille_2d(rows=10, length_factor=2),
          "lid_driven_cavity_2d": lambda: scenarios.lid_driven_cavity_2d(dx=0.04)}[name]()
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
