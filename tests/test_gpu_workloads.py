"""GPU: the benchmark workloads at their own sizes.

- 3D dam break, 1,013,040 fluid + 645,804 wall particles (examples/Cruchaga3D/
  Cruchaga3D.jl:69-164): the neighbour pair sets -- the cell-scan export and
  the pair passes' own decisions in both arithmetic forms -- bit-exact against
  the oracle's rule, at upload and after 31 steps (which cross one Shepard
  step); the fields after 31 steps against the FP64 oracle, FP64 within 1e-9
  and FP32_BULK within 1e-4 relative L2.
- The 256-scene 2D ensemble (LidDrivenCavity.jl:10-158, PoiseuilleFlow2D.jl:
  28-185) batched with graphs: 16 seeded members bit-identical to the same
  scenes stepped alone, and FP64 members of each (scene kind, kernel) within
  1e-9 of the oracle.
- The serving stream at its own sizes (2D dam breaks 5k-200k, 3D Poiseuille
  pipes 50k-2M with one of 1.24M particles, 2D thermal cavities 10k-100k)
  under a 200 GB budget: every result bit-identical to the request run alone.
"""
import os
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.serving import Server
from ethersph.solve import Batch
from oracle import oracle as O

from helpers import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

THREADS = max(1, min(16, os.cpu_count() or 1))
FIELDS = (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_C)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def keys(p):
    return (p[:, 0].astype(np.int64) << 32) | p[:, 1].astype(np.int64)


def directed(p):
    """Unordered pairs (i < j) as the sorted keys of both orders."""
    i = p[:, 0].astype(np.int64)
    j = p[:, 1].astype(np.int64)
    k = np.concatenate([(i << 32) | j, (j << 32) | i])
    k.sort()
    return k


@pytest.fixture(scope="module")
def dam3d():
    sc = scenarios.dam_break_3d()
    assert (sc.n_fluid, sc.n_wall) == (1013040, 645804)
    return sc


def wall_x(sc):
    return np.concatenate([w.x_vec_ for w in sc.walls])


def check_pairs(s, x, wx, R, label):
    """The device's cell-scan export and its walks' decisions against the
    oracle's rule applied to the same binary64 positions."""
    ff = O.neighbor_list(x, R, nthreads=THREADS)
    fw = O.neighbor_list(x, R, wx, nthreads=THREADS)
    say(label, "oracle pairs FF", len(ff), "FW", len(fw))
    assert len(ff) > 30_000_000 and len(fw) > 100_000
    a = s.neighbours(_lib.PAIRS_FLUID_FLUID)
    assert np.array_equal(keys(a), keys(ff)), f"{label}: FF export"
    del a
    assert np.array_equal(keys(s.neighbours(_lib.PAIRS_FLUID_WALL)), keys(fw)), f"{label}: FW export"
    w = s.walk_pairs(_lib.PAIRS_FLUID_FLUID)
    assert np.array_equal(keys(w), directed(ff)), f"{label}: FF walk decisions"
    del w
    assert np.array_equal(keys(s.walk_pairs(_lib.PAIRS_FLUID_WALL)), keys(fw)), f"{label}: FW walk decisions"


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32_BULK])
def test_dam_break_3d_full_size_pairs_at_upload(ctx, dam3d, prec):
    say("dam break 3D full size: pairs at upload, prec", prec)
    s = dam3d.make_scene(prec, ctx)
    check_pairs(s, dam3d.fluid.x_vec_, wall_x(dam3d), dam3d.kernel.influence_radius_, f"upload prec {prec}")
    s.close()


@pytest.fixture(scope="module")
def dam3d_oracle_31(dam3d):
    say("oracle: 31 steps of the full 3D dam break on", THREADS, "threads")
    o = O.from_scenario(dam3d)
    o.step(31, THREADS)
    ref = {f: o.get(f) for f in (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P)}
    o.close()
    return ref


@pytest.mark.parametrize("prec,bar", [(_lib.FP64, 1e-9), (_lib.FP32_BULK, 1e-4)])
def test_dam_break_3d_full_size_31_steps_against_oracle(ctx, dam3d, dam3d_oracle_31, prec, bar):
    """31 steps (one Shepard step at step 30) of the full scene: positions
    compared by their displacement from the start, the other fields as they
    are; then the pair sets on the device's own moved positions."""
    ref = dam3d_oracle_31
    say("dam break 3D full size: 31 steps, prec", prec)
    s = dam3d.make_scene(prec, ctx)
    s.step(31)
    s.sync()
    x0 = dam3d.fluid.x_vec_
    errs = {}
    for f, r in ref.items():
        a = s.get(f)
        errs[f] = rel_l2(a - x0, r - x0) if f == _lib.FIELD_X else rel_l2(a, r)
    say("relative L2 after 31 steps", prec, errs)
    for f, e in errs.items():
        assert e < bar, (prec, f, e)
    check_pairs(s, s.get(_lib.FIELD_X), wall_x(dam3d), dam3d.kernel.influence_radius_, f"31 steps prec {prec}")
    s.close()


def test_ensemble_256_scenes(ctx):
    """The ensemble of 256 scenes stepped as one batch (graph-replayed).
    16 seeded FP32_BULK members equal the same scenes alone bit for bit;
    the smallest scene of each (kind, kernel) class is made FP64 and checked
    against the oracle."""
    scs = scenarios.ensemble_2d(n_scenes=256, seed=0, total_step=5000)
    kind = ["cavity" if sc.name.startswith("lid") else "channel" for sc in scs]
    kern = [sc.kernel.kind for sc in scs]
    fp64 = []
    for k_ in ("cavity", "channel"):
        for kk in (_lib.CUBIC_SPLINE, _lib.WENDLAND_C2):
            idx = [q for q in range(256) if kind[q] == k_ and kern[q] == kk]
            fp64.append(min(idx, key=lambda q: scs[q].n_fluid))
    rng = np.random.default_rng(3)
    pick = sorted(int(q) for q in rng.choice([q for q in range(256) if q not in fp64], 16, replace=False))
    assert {kind[q] for q in pick} == {"cavity", "channel"} and {kern[q] for q in pick} == {0, 2}
    say("ensemble: members", len(scs), "particles", sum(sc.n_fluid for sc in scs), "fp64", fp64)
    scenes = [sc.make_scene(_lib.FP64 if q in fp64 else _lib.FP32_BULK, ctx) for q, sc in enumerate(scs)]
    steps = 61
    b = Batch(scenes, graph=True)
    b.step(40)
    b.step(steps - 40)
    b.sync()
    for q in pick:
        a = scs[q].make_scene(_lib.FP32_BULK, ctx)
        a.step(steps)
        a.sync()
        m = scenes[q]
        assert a.fluid_count() == m.fluid_count(), q
        for f in FIELDS + (_lib.FIELD_ID,):
            assert np.array_equal(a.get(f), m.get(f)), (q, scs[q].name, f)
        a.close()
    for q in fp64:
        o = O.from_scenario(scs[q])
        o.step(steps, THREADS)
        m = scenes[q]
        assert m.fluid_count() == o.fluid_count(), q
        assert np.array_equal(m.get(_lib.FIELD_ID), o.get(_lib.FIELD_ID)), q
        x0 = None
        for f in (_lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_X):
            e = rel_l2(m.get(f), o.get(f))
            say("fp64 member", q, scs[q].name, scs[q].n_fluid, "field", f, e)
            assert e < 1e-9, (q, f, e)
        o.close()
    b.close()
    for s in scenes:
        s.close()


def test_serving_stream_at_workload_sizes():
    """Eight Poisson arrivals of the serving workload at its own sizes (seed
    16: a 1,236,528-particle 3D Poiseuille pipe among them) under a 200 GB
    budget, 20-60 steps each; requests are submitted in arrival order while
    the batch runs.  Each result equals the request run alone."""
    ctx = _lib.Context(0, 200_000_000_000)
    stream = scenarios.serving_stream(n_requests=8, seed=16, rate_per_s=2.0, steps=(20, 60))
    big = [sc for _, sc, _ in stream if sc.name == "poiseuille_3d" and sc.n_fluid >= 1_000_000]
    assert big and {sc.name for _, sc, _ in stream} == {"collapse_dry_2d", "poiseuille_3d", "thermal_cavity_2d"}
    say("serving stream", [(sc.name, sc.n_fluid, n) for _, sc, n in stream])
    sv = Server(ctx)
    ids = []
    for _, sc, n in stream:
        ids.append(sv.submit_scenario(sc, _lib.FP32_BULK, n))
        sv.advance(7)
    for _ in range(100):
        if all(sv.status(r)[0] in (_lib.REQ_DONE, _lib.REQ_FAILED) for r in ids):
            break
        sv.advance(13)
    st = sv.stats()
    say("server stats", st)
    assert all(sv.status(r) == (_lib.REQ_DONE, n) for r, (_, _, n) in zip(ids, stream))
    assert st["completed"] == 8 and st["p99_step_ms"] >= st["p50_step_ms"] > 0
    for rid, (_, sc, n) in zip(ids, stream):
        s = sc.make_scene(_lib.FP32_BULK, ctx)
        s.step(n)
        s.sync()
        fields = FIELDS + (_lib.FIELD_ID,) + ((_lib.FIELD_T,) if s.thermal else ())
        for f in fields:
            assert np.array_equal(sv.result(rid, f), s.get(f)), (sc.name, sc.n_fluid, f)
        s.close()
        sv.release(rid)
    sv.close()
    ctx.close()
