"""VTP output and input: mirror of DataIO/VTPIO.jl and the reading half of
DataIO/VTPToDataFrame.jl.

``writeVTP(step, t, vtp_io, particle_groups)`` writes one VTK XML PolyData
file with the reference's layout (VTPIO.jl:41-98): every particle of every
group is a vertex; point data ``Velocity`` (dim components; NaN for groups
that do not move), ``Type`` (1-based group index) and the user fields named
by ``field_names_`` (NaN where a group lacks the field); field data ``Step``,
``Time`` and ``WallTime`` (format of DataIO.jl:10).  Arrays are written as
uncompressed appended raw binary with UInt64 headers.

``AsyncVTPWriter`` is the device-side companion: a scene's state is read back
at an output step and encoded and written by a host thread while the device
advances the next steps.
"""
from __future__ import annotations

import datetime
import os
import queue
import threading
import xml.etree.ElementTree as ET

import numpy as np

from . import _lib
from .particles import LiquidParticles, WallParticles

K_POSITION_STRINGS = ["X", "Y", "Z"]
K_VELOCITY_STRING = "Velocity"
K_VELOCITY_STRINGS = ["VelocityX", "VelocityY", "VelocityZ"]
K_TYPE_STRING = "Type"
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int64): "Int64",
              np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8", np.dtype(np.uint64): "UInt64"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


class VTPIO:
    """VTPIO.jl:8-15."""

    def __init__(self, step_digit: int, file_name: str, file_suffix: str, dir_path: str, field_symbols,
                 field_names):
        self.step_digit_ = int(step_digit)
        self.file_name_ = file_name
        self.file_suffix_ = file_suffix
        self.dir_path_ = dir_path
        self.field_symbols_ = [str(s).lstrip(":") for s in field_symbols]
        self.field_names_ = list(field_names)


def vtpFileNameAtStep(step: int, vtp_io: VTPIO) -> str:
    """VTPIO.jl:17-19: step zero-padded to step_digit_ digits."""
    return os.path.join(vtp_io.dir_path_, f"{vtp_io.file_name_}{int(step):0{vtp_io.step_digit_}d}{vtp_io.file_suffix_}")


def assureDirPathExist(vtp_io: VTPIO) -> None:
    """DataIO.jl:35-40."""
    os.makedirs(vtp_io.dir_path_, exist_ok=True)


def _group_arrays(g, field_symbols):
    """(x (n, dim), v (n, dim) or None, {symbol: (n,) values or None})."""
    if isinstance(g, (LiquidParticles, WallParticles)):
        x = np.asarray(g.x_vec_, dtype=np.float64)
        moving = isinstance(g, LiquidParticles) or getattr(g, "kind", None) == _lib.WALL_STATIC_VELOCITY
        v = np.asarray(g.v_vec_, dtype=np.float64) if moving else None
        fields = {s: np.asarray(getattr(g, s), dtype=np.float64) if hasattr(g, s) else None for s in field_symbols}
        return x, v, fields
    ps = list(g)
    n = len(ps)
    dim = len(ps[0].x_vec_) if n else 0
    x = np.array([p.x_vec_ for p in ps], dtype=np.float64).reshape(n, dim)
    v = np.array([p.v_vec_ for p in ps], dtype=np.float64).reshape(n, dim) if n and hasattr(ps[0], "v_vec_") else None
    fields = {s: np.array([getattr(p, s) for p in ps], dtype=np.float64) if n and hasattr(ps[0], s) else None
              for s in field_symbols}
    return x, v, fields


def _gather(particles_group, field_symbols):
    xs, vs, types, cols = [], [], [], {s: [] for s in field_symbols}
    dim = None
    for k, g in enumerate(particles_group):
        x, v, f = _group_arrays(g, field_symbols)
        if dim is None and x.shape[0]:
            dim = x.shape[1]
        n = x.shape[0]
        xs.append(x)
        vs.append(v if v is not None else np.full_like(x, np.nan))
        types.append(np.full(n, k + 1, dtype=np.int64))
        for s in field_symbols:
            cols[s].append(f[s] if f[s] is not None else np.full(n, np.nan))
    dim = dim or 2
    X = np.concatenate([x.reshape(-1, dim) for x in xs]) if xs else np.zeros((0, dim))
    V = np.concatenate([v.reshape(-1, dim) for v in vs]) if vs else np.zeros((0, dim))
    T = np.concatenate(types) if types else np.zeros(0, dtype=np.int64)
    F = {s: (np.concatenate(c) if c else np.zeros(0)) for s, c in cols.items()}
    return dim, X, V, T, F


def _wall_time() -> str:
    # kWallTimeFormat "yyyy/mm/dd__HH:MM:SS.SS" (DataIO.jl:10)
    now = datetime.datetime.now()
    return now.strftime("%Y/%m/%d__%H:%M:%S.") + f"{now.microsecond // 10000:02d}"


def _encode(path, step, t, wall_time, dim, X, V, T, F, field_names, field_symbols) -> None:
    n = X.shape[0]
    P3 = np.zeros((n, 3))
    P3[:, :dim] = X
    blobs = []
    offset = [0]

    def arr(parent, tag_attrs, a):
        a = np.ascontiguousarray(a)
        el = ET.SubElement(parent, "DataArray", dict(tag_attrs, type=_VTK_TYPES[a.dtype], format="appended",
                                                          offset=str(offset[0])))
        raw = a.tobytes()
        blobs.append(np.uint64(len(raw)).tobytes() + raw)
        offset[0] += 8 + len(raw)
        return el

    root = ET.Element("VTKFile", type="PolyData", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    poly = ET.SubElement(root, "PolyData")
    fd = ET.SubElement(poly, "FieldData")
    arr(fd, {"Name": "Step", "NumberOfTuples": "1", "NumberOfComponents": "1"}, np.array([step], dtype=np.int64))
    arr(fd, {"Name": "Time", "NumberOfTuples": "1", "NumberOfComponents": "1"}, np.array([t], dtype=np.float64))
    wt = np.frombuffer(wall_time.encode() + b"\0", dtype=np.uint8)
    el = ET.SubElement(fd, "DataArray", {"type": "String", "Name": "WallTime", "NumberOfTuples": "1",
                                         "format": "ascii"})
    el.text = " ".join(str(int(c)) for c in wt)
    piece = ET.SubElement(poly, "Piece", NumberOfPoints=str(n), NumberOfVerts=str(n), NumberOfLines="0",
                          NumberOfStrips="0", NumberOfPolys="0")
    pts = ET.SubElement(piece, "Points")
    arr(pts, {"Name": "Points", "NumberOfComponents": "3"}, P3)
    verts = ET.SubElement(piece, "Verts")
    arr(verts, {"Name": "connectivity"}, np.arange(n, dtype=np.int64))
    arr(verts, {"Name": "offsets"}, np.arange(1, n + 1, dtype=np.int64))
    pd = ET.SubElement(piece, "PointData")
    arr(pd, {"Name": K_VELOCITY_STRING, "NumberOfComponents": str(dim)}, V)
    arr(pd, {"Name": K_TYPE_STRING, "NumberOfComponents": "1"}, T)
    for s, name in zip(field_symbols, field_names):
        arr(pd, {"Name": name, "NumberOfComponents": "1"}, F[s])
    ET.SubElement(root, "AppendedData", encoding="raw").text = "_"
    head = ET.tostring(root, encoding="unicode")
    # the appended block follows the "_" marker verbatim
    head, tail = head.split('<AppendedData encoding="raw">_</AppendedData>')
    tmp = path + ".part"
    with open(tmp, "wb") as fh:
        fh.write(b'<?xml version="1.0"?>\n')
        fh.write(head.encode())
        fh.write(b'<AppendedData encoding="raw">_')
        for b in blobs:
            fh.write(b)
        fh.write(b"</AppendedData>")
        fh.write(tail.encode())
    os.replace(tmp, path)


def writeVTP(step: int, t: float, vtp_io: VTPIO, particles_group) -> str:
    """VTPIO.jl:41-98.  Returns the file written."""
    dim, X, V, T, F = _gather(particles_group, vtp_io.field_symbols_)
    path = vtpFileNameAtStep(step, vtp_io)
    _encode(path, int(step), float(t), _wall_time(), dim, X, V, T, F, vtp_io.field_names_, vtp_io.field_symbols_)
    return path


def readVTP(path: str) -> dict:
    """Reads a file written by writeVTP: {"points": (n, 3), "point_data":
    {name: array}, "field_data": {name: value}} (the data VTPIODataFrameView
    holds, VTPToDataFrame.jl:8-57)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    marker = b'<AppendedData encoding="raw">_'
    i = raw.index(marker)
    j = raw.rindex(b"</AppendedData>")
    blob = raw[i + len(marker):j]
    root = ET.fromstring(raw[:i] + b'<AppendedData encoding="raw"/>' + raw[j + len(b"</AppendedData>"):])

    def data(el):
        if el.get("format") == "ascii":
            vals = [int(v) for v in el.text.split()]
            return bytes(vals).rstrip(b"\0").decode() if el.get("type") == "String" else np.array(vals)
        off = int(el.get("offset"))
        nbytes = int(np.frombuffer(blob[off:off + 8], dtype=np.uint64)[0])
        a = np.frombuffer(blob[off + 8:off + 8 + nbytes], dtype=_NP_TYPES[el.get("type")])
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    poly = root.find("PolyData")
    out = {"field_data": {}, "point_data": {}}
    for el in poly.find("FieldData"):
        v = data(el)
        out["field_data"][el.get("Name")] = v if isinstance(v, str) else v[0]
    piece = poly.find("Piece")
    out["points"] = data(piece.find("Points")[0])
    for el in piece.find("PointData"):
        out["point_data"][el.get("Name")] = data(el)
    return out


class AsyncVTPWriter:
    """Writes VTP files on a host thread so that the device can advance while
    the previous output is encoded and written.  At most ``depth`` outputs
    wait in memory; ``close`` waits for the last file."""

    def __init__(self, vtp_io: VTPIO, depth: int = 2):
        self.vtp_io = vtp_io
        assureDirPathExist(vtp_io)
        self.q: queue.Queue = queue.Queue(maxsize=depth)
        self.errors: list = []
        self.written: list = []
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def _run(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            try:
                self.written.append(writeVTP(*item))
            except Exception as e:  # reported by close()
                self.errors.append(e)

    def submit(self, step: int, t: float, particles_group) -> None:
        """particles_group must not be modified afterwards (pass copies)."""
        self.q.put((step, t, self.vtp_io, particles_group))

    def close(self) -> list:
        self.q.put(None)
        self.thread.join()
        if self.errors:
            raise self.errors[0]
        return self.written
