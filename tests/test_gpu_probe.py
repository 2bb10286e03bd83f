"""GPU: SPH probe interpolation (kernelValueInterpolation,
DataIO/VTPToDataFrame.jl:77-142) on the resident state against the oracle:
identical inputs within 1e-13 relative (FP64 sums in another order), and
after N steps within the N-step field tolerance."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from oracle import oracle as O

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def probe_points(sc, n, seed):
    rng = np.random.default_rng(seed)
    X = sc.fluid.x_vec_
    lo, hi = X.min(axis=0), X.max(axis=0)
    span = hi - lo
    return lo - 0.2 * span + rng.random((n, sc.dim)) * 1.4 * span  # some outside every support


@pytest.mark.parametrize("name", ["collapse_dry_2d", "dam_break_3d", "thermal_cavity_2d", "poiseuille_2d"])
def test_probe_matches_oracle(ctx, name):
    sc = {"collapse_dry_2d": lambda: scenarios.collapse_dry_2d(dx=0.04),
          "dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.06),
          "thermal_cavity_2d": lambda: scenarios.thermal_cavity_2d(dx=0.03, seed=4),
          "poiseuille_2d": lambda: scenarios.poiseuille_2d(rows=10, length_factor=2)}[name]()
    say("probe", name)
    dev = sc.make_scene(_lib.FP64, ctx)
    ref = O.from_scenario(sc)
    pts = probe_points(sc, 3000, 7)
    fields = [_lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_V] + ([_lib.FIELD_T] if dev.thermal else [])
    for f in fields:
        a, b = dev.probe(pts, f, none_value=-1.0), ref.probe(pts, f, none_value=-1.0)
        assert np.array_equal(a == -1.0, b == -1.0)
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=0)
    dev.step(60)
    ref.step(60, 4)
    dev.sync()
    for f in fields:
        a, b = dev.probe(pts, f, none_value=0.0), ref.probe(pts, f, none_value=0.0)
        assert rel_l2(a, b) < 1e-9, f
    dev.close()
