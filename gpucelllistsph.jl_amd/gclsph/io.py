"""Output of an SPH problem: DataFrame / CSV (src/auxillary.jl:18-70) and VTK
PolyData files with a ParaView time collection (src/writevtk.jl:1-43,
src/sphproblem.jl:444-490).

The data come from the problem's getters (one device-to-host copy per field);
the files are written on the host.  VTK files use the XML PolyData format with
zlib-compressed inline binary arrays, the layout the reference produces with
`vtk_grid(...; compress = true, append = false)`: points (3 components, zero
z in 2D), one point-data array per requested variable (vectors with one
component per dimension), and empty vertex / polygon cell lists.
"""
from __future__ import annotations

import base64
import os
import xml.etree.ElementTree as ET
import zlib
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

_AXES = ("x", "y", "z")

# makedf / writevtk variable names (src/writevtk.jl:23, src/auxillary.jl:18)
VTKVARS = ("Density", "Pressure", "Type", "Acceleration", "Velocity", "∑W", "∑∇W", "DPC")


def _getters():
    from . import api
    return {
        "points": api.get_points,
        "Density": api.get_density,
        "Pressure": api.get_pressure,
        "Type": api.get_ptype,
        "Acceleration": api.get_acceleration,
        "Velocity": api.get_velocity,
        "∑W": api.get_sumw,
        "∑∇W": api.get_sumgradw,
        "DPC": api.get_dpccorr,
    }


def makedf(prob, vtkvars: Sequence[str] = ("Density", "Acceleration", "Velocity")):
    """makedf(prob; vtkvars) — src/auxillary.jl:18-59: a pandas DataFrame with
    columns x, y[, z], then Density, Pressure, Acceleration_<axis>,
    Velocity_<axis>, ∑W, ∑∇W_<axis>, DPC_<axis> for the requested variables,
    in that order (Type is not a makedf column in the reference)."""
    import pandas as pd
    g = _getters()
    pts = g["points"](prob)
    dim = len(pts)
    cols: List[Tuple[str, np.ndarray]] = [(_AXES[i], np.asarray(pts[i])) for i in range(dim)]
    if "Density" in vtkvars:
        cols.append(("Density", np.asarray(g["Density"](prob))))
    if "Pressure" in vtkvars:
        cols.append(("Pressure", np.asarray(g["Pressure"](prob))))
    for name in ("Acceleration", "Velocity"):
        if name in vtkvars:
            arrs = g[name](prob)
            cols += [(f"{name}_{_AXES[i]}", np.asarray(arrs[i])) for i in range(dim)]
    if "∑W" in vtkvars:
        cols.append(("∑W", np.asarray(g["∑W"](prob))))
    for name in ("∑∇W", "DPC"):
        if name in vtkvars:
            arrs = g[name](prob)
            cols += [(f"{name}_{_AXES[i]}", np.asarray(arrs[i])) for i in range(dim)]
    return pd.DataFrame(dict(cols))


def writecsv(prob, path: str, vtkvars: Sequence[str] = ("Density", "Acceleration", "Velocity")):
    """writecsv(prob, path; vtkvars) — src/auxillary.jl:66-70: the makedf table
    as comma-separated text with a header row.  The directory must exist."""
    dirn = os.path.dirname(path)
    if not os.path.isdir(dirn if dirn else "."):
        raise ValueError("Wrong path.")
    makedf(prob, vtkvars).to_csv(path, index=False, float_format="%.17g")


# ---------------------------------------------------------------- VTK writer
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int32): "Int32",
              np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}


def _encode(a: np.ndarray) -> str:
    """VTK XML inline binary with zlib compression, one block: base64(header)
    followed by base64(compressed data); header = UInt64 [blocks, block size,
    last block size, compressed size]."""
    raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
    comp = zlib.compress(raw)
    header = np.array([1, len(raw), len(raw), len(comp)], dtype="<u8").tobytes()
    return base64.b64encode(header).decode("ascii") + base64.b64encode(comp).decode("ascii")


def _data_array(parent: ET.Element, name: str, a: np.ndarray, ncomp: int = 1) -> None:
    el = ET.SubElement(parent, "DataArray", type=_VTK_TYPES[a.dtype], Name=name, format="binary")
    if ncomp != 1:
        el.set("NumberOfComponents", str(ncomp))
    el.text = _encode(a)


def _empty_cells(parent: ET.Element, tag: str) -> None:
    cells = ET.SubElement(parent, tag)
    for nm in ("connectivity", "offsets"):
        ET.SubElement(cells, "DataArray", type="Int64", Name=nm, format="ascii").text = ""


def create_vtp_file(filename: str, x: Sequence[np.ndarray], expdict: Dict[str, object], pvd=None,
                    time: Optional[float] = None) -> str:
    """create_vtp_file (src/writevtk.jl:1-13): writes `filename`.vtp (the
    extension is added when missing) and registers it in `pvd` at `time`.
    Scalars are arrays of length n; vectors are sequences of D arrays."""
    path = filename if filename.endswith(".vtp") else filename + ".vtp"
    n = len(x[0])
    pts = np.zeros((n, 3), dtype=np.float64)
    for k in range(len(x)):
        pts[:, k] = np.asarray(x[k], dtype=np.float64)
    root = ET.Element("VTKFile", type="PolyData", version="1.0", byte_order="LittleEndian", header_type="UInt64",
                      compressor="vtkZLibDataCompressor")
    piece = ET.SubElement(ET.SubElement(root, "PolyData"), "Piece", NumberOfPoints=str(n), NumberOfVerts="0",
                          NumberOfLines="0", NumberOfStrips="0", NumberOfPolys="0")
    _data_array(ET.SubElement(piece, "Points"), "Points", pts.reshape(-1), 3)
    pdata = ET.SubElement(piece, "PointData")
    for name, v in expdict.items():
        if isinstance(v, (tuple, list)):
            comps = [np.asarray(c) for c in v]
            arr = np.stack(comps, axis=1)
            _data_array(pdata, name, arr.reshape(-1), arr.shape[1])
        else:
            arr = np.asarray(v)
            if arr.dtype not in _VTK_TYPES:
                arr = arr.astype(np.float64)
            _data_array(pdata, name, arr)
    _empty_cells(piece, "Verts")
    _empty_cells(piece, "Polys")
    ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=True)
    if pvd is not None and time is not None:
        pvd[time] = path
    return path


def collect_vtk(prob, vtkvars: Iterable[str], cpupoints=None):
    """The host copies writevtk needs: (points, {variable: array or tuple of
    arrays}, simulated time).  Only this step touches the device; the file can
    then be written from another thread while the problem steps on."""
    g = _getters()
    if cpupoints is None:
        cpupoints = g["points"](prob)
    expdict: Dict[str, object] = {}
    for name in VTKVARS:
        if name in vtkvars:
            val = g[name](prob)
            expdict[name] = tuple(val) if isinstance(val, tuple) else val
    return cpupoints, expdict, prob.etime


def writevtk(prob, filename: str, vtkvars: Iterable[str], cpupoints=None, pvd=None, writetime: bool = False) -> str:
    """writevtk(prob, filename, vtkvars, cpupoints; pvd, writetime) —
    src/writevtk.jl:29-43.  vtkvars ⊂ ("Density", "Pressure", "Type",
    "Acceleration", "Velocity", "∑W", "∑∇W", "DPC"); with writetime the file
    joins `pvd` at the problem's simulated time."""
    pts, expdict, etime = collect_vtk(prob, vtkvars, cpupoints)
    return create_vtp_file(filename, pts, expdict, pvd=pvd, time=etime if writetime else None)


class VtkWriter:
    """Writes .vtp files on one background thread, in submission order, so
    that compression and disk output overlap the next batch of steps (the
    device copies are taken on the caller's thread by collect_vtk; zlib and
    file output release the interpreter lock).  close() waits for every file
    and re-raises the first error."""

    def __init__(self):
        from concurrent.futures import ThreadPoolExecutor
        self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="vtk-writer")
        self._futures = []

    def submit(self, prob, filename: str, vtkvars: Iterable[str], pvd=None, writetime: bool = False):
        pts, expdict, etime = collect_vtk(prob, vtkvars)
        self._futures.append(self._pool.submit(create_vtp_file, filename, pts, expdict, pvd,
                                               etime if writetime else None))

    def close(self) -> List[str]:
        try:
            return [f.result() for f in self._futures]
        finally:
            self._pool.shutdown(wait=True)
            self._futures = []


class paraview_collection:  # noqa: N801  (the reference's name, WriteVTK.paraview_collection)
    """A ParaView time collection (.pvd): `pvd[time] = vtp_path` adds a data
    set, `vtk_save(pvd)` (or `save()`) writes the file."""

    def __init__(self, filename: str):
        self.path = filename if filename.endswith(".pvd") else filename + ".pvd"
        self.entries: List[Tuple[float, str]] = []

    def __setitem__(self, time: float, vtp_path: str) -> None:
        self.entries.append((float(time), vtp_path))

    def save(self) -> str:
        root = ET.Element("VTKFile", type="Collection", version="1.0", byte_order="LittleEndian")
        coll = ET.SubElement(root, "Collection")
        base = os.path.dirname(os.path.abspath(self.path))
        for t, f in self.entries:
            rel = os.path.relpath(os.path.abspath(f), base)
            ET.SubElement(coll, "DataSet", timestep=repr(t), part="0", file=rel)
        ET.ElementTree(root).write(self.path, encoding="utf-8", xml_declaration=True)
        return self.path


def vtk_save(pvd: paraview_collection) -> str:
    return pvd.save()


# ---------------------------------------------------------------- VTK reader
def _decode(text: str, vtk_type: str) -> np.ndarray:
    dt = {v: k for k, v in _VTK_TYPES.items()}[vtk_type].newbyteorder("<")
    text = "".join(text.split())
    hlen = len(base64.b64encode(bytes(32)))  # 4 UInt64 words
    header = np.frombuffer(base64.b64decode(text[:hlen]), dtype="<u8")
    nblocks = int(header[0])
    if nblocks == 0:
        return np.zeros(0, dtype=dt)
    if nblocks != 1:
        raise ValueError("only single-block arrays are read")
    data = zlib.decompress(base64.b64decode(text[hlen:]))
    return np.frombuffer(data, dtype=dt).copy()


def read_vtp(path: str) -> Dict[str, np.ndarray]:
    """Reads a .vtp written by create_vtp_file: {"Points": (n, 3), name: (n,)
    or (n, components)}."""
    root = ET.parse(path).getroot()
    piece = root.find("PolyData/Piece")
    out: Dict[str, np.ndarray] = {}
    for el in list(piece.find("Points")) + list(piece.find("PointData")):
        a = _decode(el.text or "", el.get("type"))
        nc = int(el.get("NumberOfComponents", "1"))
        out[el.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    return out
