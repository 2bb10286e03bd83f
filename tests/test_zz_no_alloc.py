"""GPU: the allocation-free probe entry point (esph_scene_probe_noalloc)
gives the same values as esph_scene_probe, bit for bit, on two calls in a
row -- the second with fewer points, so values left in the shared staging
array by the first call must not leak into it -- and rejects more points
than its staging array holds."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["collapse_dry_2d", "dam_break_3d", "thermal_cavity_2d"])
def test_probe_noalloc_equals_probe(name):
    ctx = _lib.Context(0)
    sc = {"collapse_dry_2d": lambda: scenarios.collapse_dry_2d(dx=0.04),
          "dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.05),
          "thermal_cavity_2d": lambda: scenarios.thermal_cavity_2d(dx=0.02, seed=3)}[name]()
    say("probe_noalloc", name, sc.n_fluid)
    s = sc.make_scene(_lib.FP64, ctx)
    s.step(7)
    s.sync()
    X = sc.fluid.x_vec_
    lo, hi = X.min(0) - 0.05, X.max(0) + 0.05
    rng = np.random.default_rng(4)
    fields = [_lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P] + ([_lib.FIELD_T] if s.thermal else [])
    for npts in (5000, 37):  # the second call smaller than the first
        pts = rng.uniform(lo, hi, (npts, sc.dim))
        for f in fields:
            a = s.probe_noalloc(pts, f, -7.0)
            b = s.probe(pts, f, -7.0)
            assert a.shape == b.shape and np.array_equal(a, b), (name, npts, f)
            assert np.any(a != -7.0)
    with pytest.raises(_lib.EsphError):
        s.probe_noalloc(np.zeros((65537, sc.dim)), _lib.FIELD_RHO)
    s.close()
    ctx.close()
