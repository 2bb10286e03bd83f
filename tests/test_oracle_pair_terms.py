"""Pins the oracle's fluid-wall terms, the static-velocity relative
velocity, Boussinesq buoyancy with T^{n+1} and the coincident-particle
gradient guard against closed forms of single pairs (tests/pair_cases.py
derives each expected value from the reference's formulas)."""
import numpy as np
import pytest

from ethersph import _lib
from oracle import oracle as O

import pair_cases as PC

CASES = PC.all_cases()


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_oracle_single_pair_closed_form(case):
    o = O.from_scenario(case.sc)
    o.step(case.steps)
    PC.check_fields(o.get, case)


def test_closed_forms_detect_planted_mistakes():
    """The checks above separate the reference's terms from plausible
    mistakes: each planted variant of an expected value fails check_fields."""
    c = PC.fw_pressure(False)
    o = O.from_scenario(c.sc)
    o.step(1)
    v = o.get(_lib.FIELD_V)
    # m_j for m_i is invisible with one fluid particle, so compare the FW
    # pressure against the variant without the 1% tensile correction
    rv = c.sc.fluid.x_vec_[0] - c.sc.walls[0].x_vec_[0]
    r = float(np.linalg.norm(rv))
    W, dW = PC.wendland_c2_2d(r)
    p = 10.0 ** 2 * 1000.0 * (1010.0 / 1000.0 - 1.0) + 5.0
    P = p / 1010.0 ** 2 + 5.0 / 1000.0 ** 2
    dv_plain = -0.11 * P * dW * rv / r
    assert abs(v[0, 1] - dv_plain[1] * PC.DT) > 1e-6 * abs(v[0, 1])
    # buoyancy from T^n (= t_0) would leave the particle at rest
    c = PC.fw_thermostat_and_buoyancy(False)
    o = O.from_scenario(c.sc)
    o.step(1)
    assert abs(o.get(_lib.FIELD_V)[0, 1]) > 0.0
    # the model's cp, not the particle's: the two give different T^{n+1}
    T1 = o.get(_lib.FIELD_T)[0]
    dT_model = T1 - 0.8
    assert abs(dT_model * 1.7 / 3.1 - dT_model) > 1e-3 * abs(dT_model)
    # the moving wall's normal velocity enters the wall force
    a = O.from_scenario(PC.static_velocity_wall_force(True).sc)
    b = O.from_scenario(PC.static_velocity_wall_force(False).sc)
    a.step(1)
    b.step(1)
    assert a.get(_lib.FIELD_V)[0, 1] > 1.5 * b.get(_lib.FIELD_V)[0, 1] > 0.0
