"""CPU: the oracle's inlet/outlet solver (FluidInletOutletWallSolve.jl)
against the migration rules replayed on Python lists, the buffer velocity
profile against the example's formula, and a physical check of the 2D
channel against the analytic start-up flow (PoiseuilleFlow2D.jl:28-35)."""
import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import BufferProfile
from oracle import oracle as O

from helpers import free_io_scene, migrate_by_lists


@pytest.mark.parametrize("steps", [1, 11, 21, 22, 50, 60, 75, 130, 160])
def test_migration_matches_push_and_deleteat(steps):
    sc = free_io_scene()
    o = O.from_scenario(sc)
    o.step(steps)
    inlet, outlet = sc.buffers
    fl, il, ol = migrate_by_lists(sc.fluid, inlet, outlet, steps, sc.td.dt_, sc.planes)
    for group, want in ((_lib.GROUP_FLUID, fl), (_lib.GROUP_INLET, il), (_lib.GROUP_OUTLET, ol)):
        assert o.group_count(group) == len(want), (group, steps)
        if not want:
            continue
        assert np.array_equal(o.get(_lib.FIELD_ID, group), [p[2] for p in want])
        assert np.array_equal(o.get(_lib.FIELD_X, group), np.array([p[0] for p in want]))
        assert np.array_equal(o.get(_lib.FIELD_V, group), np.array([p[1] for p in want]))


def test_migration_events_happen():
    """The scene is built so that every rule fires within 160 steps."""
    sc = free_io_scene()
    inlet, outlet = sc.buffers
    fl, il, ol = migrate_by_lists(sc.fluid, inlet, outlet, 60, sc.td.dt_, sc.planes)
    assert max(p[2] for p in fl) >= len(sc.fluid) + 2   # inlet copies joined the fluid
    assert sum(p[2] >= len(outlet) for p in ol) == 3     # fluid moved into the outlet
    assert 0 not in [p[2] for p in ol]                   # an outlet particle expired
    fl, il, ol = migrate_by_lists(sc.fluid, inlet, outlet, 160, sc.td.dt_, sc.planes)
    assert len(ol) == 1                                  # the leavers expired too
    assert max(p[2] for p in fl) == len(sc.fluid) + 5    # every inlet particle entered twice but one


def test_buffer_profile_time_is_step_times_dt():
    """user_defined_function(t = step * dt) after step 1 (steps count from 1):
    inlet v_x = uxTheory(y, dt); outlet v_x too and v_y = 0 (2D example)."""
    f, nu, l = 8e-5, 1e-6, 1e-2
    prof = BufferProfile(_lib.PROFILE_CHANNEL_2D, f, nu, l, (0.0, 0.0, 0.0), terms=10, outlet_profile=True)
    sc = free_io_scene(profile=prof)
    # put the buffers inside the channel height
    for b in sc.buffers:
        b.x_vec_[:, 1] = np.linspace(0.001, 0.009, len(b))
        b.v_vec_[:, 1] = 0.3
    o = O.from_scenario(sc)
    o.step(1)
    dt = sc.td.dt_
    for g in (_lib.GROUP_INLET, _lib.GROUP_OUTLET):
        x = o.get(_lib.FIELD_X, g)
        v = o.get(_lib.FIELD_V, g)
        want = scenarios._channel_ux(x[:, 1], 1 * dt, f, nu, l, 10)
        np.testing.assert_allclose(v[:, 0], want, rtol=1e-12, atol=1e-18)
        if g == _lib.GROUP_OUTLET:
            assert np.all(v[:, 1] == 0.0)
        else:
            assert np.all(v[:, 1] == 0.3)
    # 3D pipe: steady profile on the inlet; the outlet keeps v_x
    p3 = scenarios.poiseuille_3d(cells=8, length_factor=1)
    o3 = O.from_scenario(p3)
    vx_out0 = p3.buffers[1].v_vec_[:, 0]
    o3.step(1)
    x = o3.get(_lib.FIELD_X, _lib.GROUP_INLET)
    v = o3.get(_lib.FIELD_V, _lib.GROUP_INLET)
    r0 = 0.5e-2
    np.testing.assert_allclose(v[:, 0], f / 4 / nu * (r0 ** 2 - x[:, 1] ** 2 - x[:, 2] ** 2), rtol=1e-12)
    vo = o3.get(_lib.FIELD_V, _lib.GROUP_OUTLET)
    assert np.all(vo[:, 1:] == 0.0)
    assert set(np.round(vo[:, 0], 15)) <= set(np.round(vx_out0, 15))


def test_channel_tracks_the_startup_solution():
    """16 rows across a channel 2 l long, 1,500 steps (t = 28 s, nu t / l^2 =
    0.28): the fluid's axial velocity in the middle half follows the analytic
    start-up profile to 30 % of u_max pointwise and 10 % on average; the
    cross velocity stays below 10 % of u_max and is mirror-symmetric."""
    sc = scenarios.poiseuille_2d(rows=16, length_factor=2)
    o = O.from_scenario(sc)
    o.step(1500, 4)
    t = 1500 * sc.td.dt_
    x = o.get(_lib.FIELD_X)
    v = o.get(_lib.FIELD_V)
    l, f, nu = 1e-2, 8e-5, 1e-6
    umax = f / 8 / nu * l * l
    sel = (x[:, 0] > 0.5 * l) & (x[:, 0] < 1.5 * l)
    ua = scenarios._channel_ux(x[sel, 1], t, f, nu, l, 10)
    err = np.abs(v[sel, 0] - ua) / umax
    assert err.max() < 0.3 and err.mean() < 0.1, (err.max(), err.mean())
    assert np.abs(v[sel, 1]).max() < 0.1 * umax
    # mirror symmetry about the centre line: sum of v_y vanishes
    assert abs(v[sel, 1].sum()) < 1e-3 * umax * sel.sum()
    # inflow and outflow balance: the fluid count stays within 10 %
    assert abs(o.group_count(_lib.GROUP_FLUID) - sc.n_fluid) < 0.1 * sc.n_fluid


def test_pipe_scenario_shape():
    """createCylinderRingParticles (PoiseuilleFlow3D.jl:60-93): ring counts
    round(2 pi r / dr) per radius, masses len_dx * dr_theta * dr_r * rho."""
    sc = scenarios.poiseuille_3d(cells=40)
    per_slice = sum(max(1, round(2 * np.pi * (i - 0.5))) for i in range(1, 21))
    assert sc.n_fluid == 80 * per_slice
    assert len(sc.buffers[0]) == len(sc.buffers[1]) == 3 * per_slice
    assert sc.solver == _lib.SOLVER_FLUID_INLET_OUTLET_WALL
    r = np.sqrt(sc.fluid.x_vec_[:, 1] ** 2 + sc.fluid.x_vec_[:, 2] ** 2)
    assert r.max() < 0.5e-2
    wr = np.sqrt(sc.walls[0].x_vec_[:, 1] ** 2 + sc.walls[0].x_vec_[:, 2] ** 2)
    assert wr.min() > 0.5e-2
    np.testing.assert_allclose(np.einsum("ij,ij->i", sc.walls[0].normal_vec_, sc.walls[0].normal_vec_), 1.0)
