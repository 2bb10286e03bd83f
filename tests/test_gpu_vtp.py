"""GPU: VTP output of a device run (DataIO/VTPIO.jl:41-98, the output steps
of Solve/FluidWallSolve.jl:95,136).  solve(..., vtp_io) steps on the device,
reads the state back at each output step and writes the frames from a host
thread; each frame holds the oracle's state at its step (FP64, within 1e-9
relative L2), with the reference's field names, group types, NaN fields for
walls, and Step/Time field data."""
import os
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import solve
from ethersph.vtp import VTPIO, readVTP, vtpFileNameAtStep
from oracle import oracle as O

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["collapse_dry_2d", "dam_break_3d"])
def test_vtp_frames_of_a_device_run_hold_the_oracle_state(ctx, tmp_path, name):
    sc = {"collapse_dry_2d": lambda: scenarios.collapse_dry_2d(dx=0.05),
          "dam_break_3d": lambda: scenarios.dam_break_3d(dx=0.06)}[name]()
    sc.td.total_step_ = 45
    sc.td.output_interval_ = 20
    io = VTPIO(4, name, ".vtp", str(tmp_path), ["rho_", "p_", "mass_"], ["Density", "Pressure", "Mass"])
    fluid = sc.fluid.subset(np.arange(sc.n_fluid))
    x0 = fluid.x_vec_.copy()
    say("vtp frames", name, sc.n_fluid)
    solve(fluid, *sc.walls, sc.kernel, sc.model, sc.td, io, check=sc.check, precision="FP64", ctx=ctx)
    frames = sorted(os.listdir(tmp_path))
    assert frames == [os.path.basename(vtpFileNameAtStep(k, io)) for k in (0, 1, 2)], frames
    ref = O.from_scenario(sc)
    nw = sc.n_wall
    D = sc.dim
    done = 0
    for k, step in enumerate((0, 20, 40)):
        ref.step(step - done, 4)
        done = step
        d = readVTP(vtpFileNameAtStep(k, io))
        assert d["field_data"]["Step"] == k and np.isclose(d["field_data"]["Time"], step * sc.td.dt_)
        n = ref.fluid_count()
        pts, pd = d["points"], d["point_data"]
        assert pts.shape == (n + nw, 3)
        np.testing.assert_array_equal(pd["Type"], [1] * n + [2] * nw)
        assert np.all(np.isnan(pd["Density"][n:]))
        xr = ref.get(_lib.FIELD_X)
        if step == 0:
            np.testing.assert_array_equal(pts[:n, :D], x0)
        else:
            assert rel_l2(pts[:n, :D] - x0, xr - x0) < 1e-9
        assert rel_l2(pd["Velocity"][:n], ref.get(_lib.FIELD_V)) < 1e-9 or step == 0
        assert rel_l2(pd["Density"][:n], ref.get(_lib.FIELD_RHO)) < 1e-9
        assert rel_l2(pd["Pressure"][:n], ref.get(_lib.FIELD_P)) < 1e-9 or step == 0
    # the final state written back equals the oracle after all 45 steps
    ref.step(45 - done, 4)
    assert rel_l2(fluid.x_vec_ - x0, ref.get(_lib.FIELD_X) - x0) < 1e-9
