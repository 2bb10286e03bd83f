"""GPU: continuous batching.  Requests submitted while others run join the
running batch; each result equals the same scene run alone (bit-identical:
a particle's arithmetic does not depend on its batch mates), and admission
respects the HBM budget."""
import sys

import numpy as np
import pytest

from ethersph import _lib, scenarios
from ethersph.serving import Server

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(sc, prec, steps, ctx):
    s = sc.make_scene(prec, ctx)
    s.step(steps)
    s.sync()
    out = {f: s.get(f) for f in (_lib.FIELD_X, _lib.FIELD_V, _lib.FIELD_RHO, _lib.FIELD_P, _lib.FIELD_C)}
    if s.thermal:
        out[_lib.FIELD_T] = s.get(_lib.FIELD_T)
    s.close()
    return out


def test_requests_in_the_mix_equal_alone(ctx):
    prec = _lib.FP32_BULK
    jobs = [
        (scenarios.collapse_dry_2d(dx=0.04), 120),
        (scenarios.thermal_cavity_2d(dx=0.03, seed=2), 90),
        (scenarios.lid_driven_cavity_2d(dx=0.03), 150),
        (scenarios.dam_break_3d(dx=0.07), 60),
        (scenarios.collapse_dry_2d(dx=0.05, kernel=scenarios.WendlandC2), 75),
    ]
    sv = Server(ctx)
    ids = []
    # arrivals interleaved with steps of the running batch
    ids.append(sv.submit(jobs[0][0].desc(prec), jobs[0][0].fluid, jobs[0][0].walls, jobs[0][1]))
    ids.append(sv.submit(jobs[1][0].desc(prec), jobs[1][0].fluid, jobs[1][0].walls, jobs[1][1]))
    assert sv.advance(17) == 17
    ids.append(sv.submit(jobs[2][0].desc(prec), jobs[2][0].fluid, jobs[2][0].walls, jobs[2][1]))
    sv.advance(31)
    ids.append(sv.submit(jobs[3][0].desc(prec), jobs[3][0].fluid, jobs[3][0].walls, jobs[3][1]))
    ids.append(sv.submit(jobs[4][0].desc(prec), jobs[4][0].fluid, jobs[4][0].walls, jobs[4][1]))
    for _ in range(200):
        if all(sv.status(r)[0] == _lib.REQ_DONE for r in ids):
            break
        sv.advance(25)
    st = sv.stats()
    say("server stats", st)
    assert st["completed"] == len(jobs) and st["running"] == 0 and st["queued"] == 0
    assert st["p99_step_ms"] >= st["p50_step_ms"] > 0
    for rid, (sc, steps) in zip(ids, jobs):
        assert sv.status(rid) == (_lib.REQ_DONE, steps)
        ref = alone(sc, prec, steps, ctx)
        for f, v in ref.items():
            assert np.array_equal(sv.result(rid, f), v), (sc.name, f)
        sv.release(rid)
    sv.close()


def test_admission_waits_for_budget():
    """With a budget that holds one scene, the second request waits in the
    queue until the first finishes; a request that can never fit fails."""
    sc = scenarios.collapse_dry_2d(dx=0.04)
    probe = _lib.Context(0)
    s = sc.make_scene(_lib.FP64, probe)
    one = probe.memory_used()
    s.close()
    probe.close()
    ctx = _lib.Context(0, int(one * 1.5))
    sv = Server(ctx)
    a = sv.submit(sc.desc(_lib.FP64), sc.fluid, sc.walls, 40)
    b = sv.submit(sc.desc(_lib.FP64), sc.fluid, sc.walls, 40)
    big = scenarios.collapse_dry_2d(dx=0.02)
    c = sv.submit(big.desc(_lib.FP64), big.fluid, big.walls, 10)
    sv.advance(10)
    assert sv.status(a)[0] == _lib.REQ_RUNNING and sv.status(b)[0] == _lib.REQ_QUEUED
    for _ in range(20):
        if sv.status(b)[0] in (_lib.REQ_DONE, _lib.REQ_FAILED):
            break
        sv.advance(40)
    assert sv.status(b) == (_lib.REQ_DONE, 40)
    assert sv.status(a) == (_lib.REQ_DONE, 40)
    assert sv.status(c)[0] == _lib.REQ_FAILED
    assert np.array_equal(sv.result(a, _lib.FIELD_X), sv.result(b, _lib.FIELD_X))
    sv.close()
    ctx.close()


def test_blown_up_request_fails_without_stalling_the_batch(ctx):
    """A request stepped with a time step far beyond the viscous limit blows
    up; the server fails it (non-finite or far-escaped particles) while a
    well-posed request in the same batch still completes."""
    bad = scenarios.poiseuille_2d(rows=12, length_factor=2)
    bad.td.dt_ *= 40.0
    good = scenarios.collapse_dry_2d(dx=0.05)
    sv = Server(ctx)
    a = sv.submit_scenario(bad, _lib.FP32_BULK, 3000)
    b = sv.submit_scenario(good, _lib.FP32_BULK, 300)
    for _ in range(100):
        if all(sv.status(r)[0] in (_lib.REQ_DONE, _lib.REQ_FAILED) for r in (a, b)):
            break
        sv.advance(30)
    assert sv.status(a)[0] == _lib.REQ_FAILED
    assert sv.status(b) == (_lib.REQ_DONE, 300)
    sv.close()
