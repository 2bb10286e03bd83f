"""GPU: the force pass's two-record walk of 3D FP32_BULK scenes (each
candidate's p / rho^2 from the equation of state, evaluated on the candidate's
rho - rho0) in its three forms -- the linear equation of state (gamma 1, p0
0), gamma 7 (the sum (rho/rho0)^gamma - 1 = e (1 + u + ... + u^6) unrolled)
and any other gamma (the same sum as a loop) -- against the FP64 form, which
matches the FP64 oracle to 1e-9 (the parity tests).  A wrong exponent or a
dropped term of the sum changes every candidate's pressure and moves the
fields by far more than the bars.  Bars: relative L2 after 120 steps (four
Shepard re-initialisations), about 30x what was measured for every form
(at most x 2.5e-10, v 3.9e-7, rho 1.1e-9, p 6.8e-7)."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios

from helpers import rel_l2

pytestmark = pytest.mark.gpu

BARS = {"x": 1e-8, "v": 1e-5, "rho": 3e-8, "p": 2e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fields(sc, prec, steps, ctx):
    s = sc.make_scene(prec, ctx)
    s.step(steps)
    s.sync()
    out = {n: s.get(f) for n, f in (("x", _lib.FIELD_X), ("v", _lib.FIELD_V), ("rho", _lib.FIELD_RHO),
                                     ("p", _lib.FIELD_P))}
    assert np.array_equal(s.get(_lib.FIELD_ID), np.arange(sc.n_fluid))
    s.close()
    return out


@pytest.mark.parametrize("gamma", [1, 7, 5])
def test_eos_walk_forms_match_fp64(ctx, gamma):
    sc = scenarios.dam_break_3d(dx=0.03, gamma=gamma)
    say("EOS walk, 3D dam break dx 0.03, gamma", gamma, "fluid", sc.n_fluid, "wall", sc.n_wall)
    a = fields(sc, _lib.FP32_BULK, 120, ctx)
    b = fields(sc, _lib.FP64, 120, ctx)
    errs = {k: rel_l2(a[k], b[k]) for k in a}
    say("  FP32_BULK vs FP64 after 120 steps, rel-L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert all(errs[k] < BARS[k] for k in BARS), (gamma, errs)
