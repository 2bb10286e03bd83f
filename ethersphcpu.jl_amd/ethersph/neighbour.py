"""Neighbour search: the device replacement of CellListMap.neighborlist as
used by findNeighbours! (Neighbour/Neighbour.jl:88-128).

Rule (include/ethersph.h): i, j are neighbours iff d2 <= R*R with d2 summed in
binary64 without fused multiply-add; pairs inside one set have i < j.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int32, c_int64

import numpy as np

from . import _lib


def neighborlist(x, cutoff: float, y=None, ctx=None) -> np.ndarray:
    """Sorted (m, 2) int32 array of pairs (i, j): within x (i < j) when y is
    None, else between x (i) and y (j)."""
    ctx = ctx or _lib.default_context()
    X = _lib.f64(x)
    dim = X.shape[1]
    Y = _lib.f64(y) if y is not None else None
    ny = 0 if Y is None else Y.shape[0]
    m = c_int64()
    args = (ctx.handle, dim, float(cutoff), X.shape[0], _lib.dptr(X), ny, _lib.dptr(Y))
    ctx.check(ctx._L.esph_neighbor_list(*args, None, 0, byref(m)))
    out = np.empty((max(m.value, 1), 2), dtype=np.int32)
    ctx.check(ctx._L.esph_neighbor_list(*args, out.ctypes.data_as(ctypes.POINTER(c_int32)), out.shape[0], byref(m)))
    return out[: m.value]
