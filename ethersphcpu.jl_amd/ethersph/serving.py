"""Serving: continuous batching of simulation requests (esph_server).

Requests are queued on submit; ``advance`` admits them in arrival order
against the context's HBM budget, steps every running request together as
one batch and retires finished ones.  Results stay on the host until
released.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import _lib
from .solve import fluid_arrays


class Server:
    _close_rank = 0

    def __init__(self, ctx: _lib.Context, max_active: int = 0):
        self.ctx = ctx
        self.L = ctx._L
        h = c_void_p()
        ctx.check(self.L.esph_server_create(ctx.handle, max_active, byref(h)))
        self.handle = h
        ctx.adopt(self)
        self.meta = {}

    def submit(self, desc: _lib.SceneDesc, fluid, walls, nsteps: int, buffers: tuple | None = None) -> int:
        """Queues a request; fluid is a LiquidParticles, walls a list of
        WallParticles, buffers the (inlet, outlet) LiquidParticles of the
        inlet/outlet solver."""
        keep = []
        warr = (_lib.WallArrays * max(len(walls), 1))()
        for q, w in enumerate(walls):
            a = [_lib.f64(getattr(w, f)) for f in ("x_vec_", "normal_vec_", "gap_", "v_vec_", "t_", "kappa_")]
            keep += a
            warr[q].kind, warr[q].n = w.kind, len(w)
            warr[q].x, warr[q].normal, warr[q].gap, warr[q].v, warr[q].t, warr[q].kappa = (_lib.dptr(v) for v in a)
        thermal = getattr(fluid, "thermal", False)
        fa, k0 = fluid_arrays(fluid, thermal)
        ia = oa = None
        if buffers is not None:
            ia, k1 = fluid_arrays(buffers[0], False)
            oa, k2 = fluid_arrays(buffers[1], False)
            keep += [k1, k2]
        rid = c_int64()
        self.ctx.check(self.L.esph_server_submit(self.handle, byref(desc), len(walls), warr, byref(fa),
                                                 byref(ia) if ia is not None else None,
                                                 byref(oa) if oa is not None else None, nsteps, byref(rid)))
        self.meta[rid.value] = (desc.dim, thermal)
        return rid.value

    def submit_scenario(self, sc, precision: int, nsteps: int) -> int:
        """Queues a scenes.Scenario (its descriptor at the given precision)."""
        return self.submit(sc.desc(precision), sc.fluid, sc.walls, nsteps, getattr(sc, "buffers", None))

    def advance(self, max_steps: int) -> int:
        k = c_int64()
        self.ctx.check(self.L.esph_server_advance(self.handle, max_steps, byref(k)))
        return k.value

    def status(self, rid: int):
        st = c_int32()
        done = c_int64()
        self.ctx.check(self.L.esph_server_status(self.handle, rid, byref(st), byref(done)))
        return st.value, done.value

    def result(self, rid: int, field: int) -> np.ndarray:
        n = c_int64()
        self.ctx.check(self.L.esph_server_result(self.handle, rid, field, None, 0, byref(n)))
        out = np.empty(max(n.value, 1))
        self.ctx.check(self.L.esph_server_result(self.handle, rid, field, _lib.dptr(out), out.size, byref(n)))
        out = out[: n.value]
        dim = self.meta[rid][0]
        return out.reshape(-1, dim) if field in (_lib.FIELD_X, _lib.FIELD_V) else out

    def release(self, rid: int) -> None:
        self.ctx.check(self.L.esph_server_release(self.handle, rid))

    def stats(self) -> dict:
        v = (c_double * 8)()
        self.ctx.check(self.L.esph_server_stats(self.handle, v))
        keys = ("p50_step_ms", "p99_step_ms", "completed", "scenes_per_s", "running", "queued", "device_bytes",
                "batch_steps")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.L.esph_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
