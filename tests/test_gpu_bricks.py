"""GPU: brick walks (FLAG_BRICK) against the same scenes walked from device
memory (the default).  The brick walk stages each brick's neighbour region in LDS and
reads the candidates from there; the candidates, their order and the pair
arithmetic are those of the gather walk, so every field must be
bit-identical.  Covered: list rebuilds (sorts), Shepard steps, graph replay,
a batch mixing brick and gather scenes, and the full 1,013,040-particle 3D
dam break."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.solve import Batch, Scene

pytestmark = pytest.mark.gpu

FIELDS = (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_C)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# Both contexts walk each particle's wall candidates in its own lane, as
# brick walks always do (the shared wall walk sums a wavefront's wall terms
# in an order that depends on the wavefront's particles, which differ
# between the two walks).
@pytest.fixture(scope="module")
def brick_ctx():
    c = _lib.Context(0)
    c.set_brick_walk(True)
    c.set_wall_lanes(True)
    yield c
    c.close()


@pytest.fixture(scope="module")
def gather_ctx():
    c = _lib.Context(0)
    c.set_brick_walk(False)
    c.set_wall_lanes(True)
    yield c
    c.close()


def scene_with_skin(sc, ctx, skin):
    d = sc.desc(_lib.FP32_BULK)
    d.verlet_skin = skin
    return Scene(d, sc.fluid, sc.walls, ctx)


def assert_same(a, b, what):
    assert a.fluid_count() == b.fluid_count(), what
    for f in FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f"{what}: field {f} differs"


@pytest.mark.parametrize("dx,steps,graph", [(0.03, 250, False), (0.015, 96, True)])
def test_brick_walk_equals_gather_walk(brick_ctx, gather_ctx, dx, steps, graph):
    sc = scenarios.dam_break_3d(dx=dx)
    say("brick vs gather walk: 3D dam break dx", dx, "fluid", sc.n_fluid, "steps", steps, "graph", graph)
    # (candidate-list skin 0.03: the runs cross list rebuilds after the upload)
    br, ga = (scene_with_skin(sc, c, 0.03) for c in (brick_ctx, gather_ctx))
    info = br.brick_stats()
    assert info["bricks"] and info["occupied"] > 0, info
    assert not ga.brick_stats()["bricks"]
    for s in (br, ga):
        s.enable_graph(graph)
        s.step(steps)
        s.sync()
    info = br.brick_stats()
    say("  bricks", info, "sorts", br.stats()["sorts"])
    assert info["spilled"] == 0, info
    # the upload's sort and at least one rebuild after it
    assert br.stats()["sorts"] == ga.stats()["sorts"] and br.stats()["sorts"] >= 2
    assert_same(br, ga, f"dx {dx}")
    br.close()
    ga.close()


def test_batch_with_brick_and_gather_members(brick_ctx, gather_ctx):
    # one context holds both kinds: the setting applies at upload
    scs = [scenarios.dam_break_3d(dx=0.04), scenarios.dam_break_3d(dx=0.05), scenarios.collapse_dry_2d(dx=0.05)]
    say("batch of brick and gather members", [s.n_fluid for s in scs])
    ref = [scs[0].make_scene(_lib.FP32_BULK, gather_ctx), scs[1].make_scene(_lib.FP32_BULK, gather_ctx),
           scs[2].make_scene(_lib.FP32_BULK, gather_ctx)]
    mix = [scs[0].make_scene(_lib.FP32_BULK, brick_ctx)]
    brick_ctx.set_brick_walk(False)
    try:
        mix.append(scs[1].make_scene(_lib.FP32_BULK, brick_ctx))
    finally:
        brick_ctx.set_brick_walk(True)
    mix.append(scs[2].make_scene(_lib.FP32_BULK, brick_ctx))
    assert [m.brick_stats()["bricks"] for m in mix] == [True, False, False]
    for s in ref:
        s.step(70)
        s.sync()
    b = Batch(mix, graph=True)
    b.step(40)
    b.step(30)
    b.sync()
    for k, (r, m) in enumerate(zip(ref, mix)):
        assert m.step_count == 70
        assert_same(r, m, f"member {k}")
    assert mix[0].brick_stats()["spilled"] == 0
    b.close()
    for s in ref + mix:
        s.close()


@pytest.mark.slow
def test_brick_walk_full_size_dam_break(brick_ctx, gather_ctx):
    sc = scenarios.dam_break_3d()
    say("brick vs gather walk: full 3D dam break, fluid", sc.n_fluid, "wall", sc.n_wall)
    br = sc.make_scene(_lib.FP32_BULK, brick_ctx)
    ga = sc.make_scene(_lib.FP32_BULK, gather_ctx)
    for s in (br, ga):
        s.enable_graph(True)
        s.step(31)
        s.sync()
    info = br.brick_stats()
    say("  bricks", info)
    assert info["bricks"] and info["spilled"] == 0, info
    assert_same(br, ga, "full size")
    br.close()
    ga.close()
