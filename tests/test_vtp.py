"""CPU: VTP output (DataIO/VTPIO.jl) -- file naming, array layout and a
write/read round trip with the reference's field names."""
import os

import numpy as np

from ethersph import _lib, scenarios
from ethersph.particles import CommonLiquidParticle, CompulsiveWallParticle
from ethersph.vtp import AsyncVTPWriter, VTPIO, readVTP, vtpFileNameAtStep, writeVTP


def test_file_name_pads_step():
    io = VTPIO(4, "collapse_dry", ".vtp", "/tmp/x", [":rho_", ":p_"], ["Density", "Pressure"])
    assert vtpFileNameAtStep(7, io) == os.path.join("/tmp/x", "collapse_dry0007.vtp")
    assert vtpFileNameAtStep(123456, io).endswith("collapse_dry123456.vtp")
    assert io.field_symbols_ == ["rho_", "p_"]


def test_round_trip_groups_types_and_nan_fields(tmp_path):
    sc = scenarios.collapse_dry_2d(dx=0.1)
    io = VTPIO(3, "cd", ".vtp", str(tmp_path), ["rho_", "p_", "mass_"], ["Density", "Pressure", "Mass"])
    path = writeVTP(12, 0.25, io, [sc.fluid, sc.walls[0]])
    assert path == vtpFileNameAtStep(12, io)
    d = readVTP(path)
    nf, nw = sc.n_fluid, sc.n_wall
    assert d["field_data"]["Step"] == 12 and d["field_data"]["Time"] == 0.25
    assert len(d["field_data"]["WallTime"]) == len("yyyy/mm/dd__HH:MM:SS.SS")
    assert d["points"].shape == (nf + nw, 3)
    np.testing.assert_array_equal(d["points"][:nf, :2], sc.fluid.x_vec_)
    assert np.all(d["points"][:, 2] == 0.0)
    pdata = d["point_data"]
    np.testing.assert_array_equal(pdata["Type"], [1] * nf + [2] * nw)
    np.testing.assert_array_equal(pdata["Velocity"][:nf], sc.fluid.v_vec_)
    assert np.all(np.isnan(pdata["Velocity"][nf:]))   # compulsive walls do not move
    np.testing.assert_array_equal(pdata["Density"][:nf], sc.fluid.rho_)
    assert np.all(np.isnan(pdata["Density"][nf:]))    # walls have no rho_
    assert list(pdata) == ["Velocity", "Type", "Density", "Pressure", "Mass"]


def test_particle_object_lists_and_async_writer(tmp_path):
    fl = [CommonLiquidParticle(3) for _ in range(5)]
    for k, p in enumerate(fl):
        p.x_vec_[:] = [k, 2 * k, 3 * k]
        p.v_vec_[:] = [1, 0, -1]
        p.rho_ = 1000.0 + k
    ws = [CompulsiveWallParticle(3) for _ in range(2)]
    io = VTPIO(2, "o", ".vtp", str(tmp_path / "sub"), [":rho_"], ["Density"])
    w = AsyncVTPWriter(io)
    for step in range(3):
        w.submit(step, 0.1 * step, [fl, ws])
    files = w.close()
    assert [os.path.basename(f) for f in files] == ["o00.vtp", "o01.vtp", "o02.vtp"]
    d = readVTP(files[-1])
    np.testing.assert_array_equal(d["points"][:5], [[k, 2 * k, 3 * k] for k in range(5)])
    np.testing.assert_array_equal(d["point_data"]["Density"][:5], [1000.0 + k for k in range(5)])
    assert d["point_data"]["Velocity"].shape == (7, 3)
