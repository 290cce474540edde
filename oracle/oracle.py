"""ctypes wrapper around the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  It mirrors the reference's
GPUCellList / SPHProblem objects (src/celllist.jl, src/sphproblem.jl) closely
enough to drive them the way the reference's examples do.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GCLSPH_ORACLE_LIB / GCLSPH_ORACLE_OMP_LIB: other builds of the same sources
# (the sanitizer builds of `make sanitize`)
LIB_PATH = os.environ.get("GCLSPH_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None


OMP_PATH = os.environ.get("GCLSPH_ORACLE_OMP_LIB") or os.path.join(HERE, "_build", "libcpuomp.so")
_omp = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or not os.path.exists(OMP_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.orc_gcl_create.restype = P
        L.orc_gcl_create.argtypes = [C.c_int, C.c_int, C.c_int64, C.POINTER(dp), dp, C.c_double, dp, dp,
                                     C.c_int64, C.c_int64, C.POINTER(C.c_int)]
        L.orc_gcl_destroy.argtypes = [P]
        L.orc_gcl_update.argtypes = [P, C.c_int]
        L.orc_gcl_partialupdate.argtypes = [P, C.c_int]
        L.orc_gcl_set_points.argtypes = [P, C.POINTER(dp)]
        L.orc_gcl_get_points.argtypes = [P, C.POINTER(dp)]
        L.orc_gcl_info.argtypes = [P, C.POINTER(C.c_int64), dp, dp, C.POINTER(C.c_int64)]
        L.orc_gcl_get_pairs.restype = C.c_int64
        L.orc_gcl_get_pairs.argtypes = [P, C.POINTER(C.c_int32), C.c_int64]
        L.orc_gcl_get_cells.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_gcl_maxpairs.restype = C.c_int64
        L.orc_gcl_maxpairs.argtypes = [P]
        L.orc_sph_create.restype = P
        L.orc_sph_create.argtypes = [P, dp, dp, C.POINTER(C.c_int32), C.POINTER(dp)]
        L.orc_sph_destroy.argtypes = [P]
        L.orc_sph_stepsolve.argtypes = [P, C.c_int64, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int64), dp]
        L.orc_sph_get_field.argtypes = [P, C.c_int, dp]
        L.orc_sph_get_scalar.restype = C.c_double
        L.orc_sph_get_scalar.argtypes = [P, C.c_int]
        L.orc_sph_set_scalar.argtypes = [P, C.c_int, C.c_double]
        for name in ("orc_wendland_W", "orc_wendland_dW"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_powfancy7th.restype = C.c_double
        L.orc_powfancy7th.argtypes = [C.c_double]
        L.orc_dpckernlim.restype = C.c_double
        L.orc_dpckernlim.argtypes = [C.c_int]
        L.orc_eos.restype = C.c_double
        L.orc_eos.argtypes = [C.c_double, C.c_int64, C.c_double, C.c_double, C.c_int32]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ptr_array(arrs):
    t = (C.POINTER(C.c_double) * len(arrs))()
    for k, a in enumerate(arrs):
        t[k] = _dptr(a)
    return t


class OracleError(RuntimeError):
    pass


class CellList:
    """GPUCellList(points, cellsize, dist; MIN, MAX, mppcell, mpairs) on the CPU."""

    def __init__(self, points: Sequence[np.ndarray], cellsize, dist: float, MIN=None, MAX=None, mppcell: int = 0,
                 mpairs: int = 0, prec: int = 64):
        L = lib()
        self.dim = len(points)
        self._pts = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
        n = self._pts[0].shape[0]
        cs = np.ascontiguousarray(np.broadcast_to(np.asarray(cellsize, dtype=np.float64), (self.dim,)))
        mn = None if MIN is None else np.ascontiguousarray(MIN, dtype=np.float64)
        mx = None if MAX is None else np.ascontiguousarray(MAX, dtype=np.float64)
        warn = C.c_int(0)
        h = L.orc_gcl_create(prec, self.dim, n, _ptr_array(self._pts), _dptr(cs), float(dist),
                             _dptr(mn) if mn is not None else None, _dptr(mx) if mx is not None else None,
                             int(mppcell), int(mpairs), C.byref(warn))
        if not h:
            raise OracleError(L.orc_last_error().decode())
        self.h = h
        self.n = n
        self.warn = warn.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_gcl_destroy(self.h)
            self.h = None

    def update(self, fillzero: bool = True):
        lib().orc_gcl_update(self.h, int(fillzero))

    def partialupdate(self, fillzero: bool = True):
        lib().orc_gcl_partialupdate(self.h, int(fillzero))

    def set_points(self, points):
        arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
        lib().orc_gcl_set_points(self.h, _ptr_array(arrs))

    def points(self) -> Tuple[np.ndarray, ...]:
        out = [np.empty(self.n) for _ in range(self.dim)]
        lib().orc_gcl_get_points(self.h, _ptr_array(out))
        return tuple(out)

    def info(self):
        grid = (C.c_int64 * 3)()
        off = np.zeros(3)
        cs = np.zeros(3)
        sc = (C.c_int64 * 4)()
        lib().orc_gcl_info(self.h, grid, _dptr(off), _dptr(cs), sc)
        return {"grid": tuple(grid[: self.dim]), "offset": tuple(off[: self.dim]), "cs": tuple(cs[: self.dim]),
                "n": sc[0], "mppcell": sc[1], "capacity": sc[2], "pairsn": sc[3]}

    @property
    def pairsn(self) -> int:
        return self.info()["pairsn"]

    def neighborlist(self) -> np.ndarray:
        cap = self.info()["capacity"]
        out = np.zeros((cap, 2), dtype=np.int32)
        lib().orc_gcl_get_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        return out

    def neighborlistview(self) -> np.ndarray:
        inf = self.info()
        if inf["pairsn"] > inf["capacity"]:
            raise OracleError("pair list overflow: pairsn > capacity")
        return self.neighborlist()[: inf["pairsn"]]

    def cells(self):
        inf = self.info()
        pc = np.zeros(self.dim * self.n, dtype=np.int32)
        ncell = int(np.prod(inf["grid"]))
        cp = np.zeros(ncell, dtype=np.int32)
        lib().orc_gcl_get_cells(self.h, pc.ctypes.data_as(C.POINTER(C.c_int32)),
                                cp.ctypes.data_as(C.POINTER(C.c_int32)))
        return pc.reshape(self.dim, self.n), cp.reshape(tuple(reversed(inf["grid"]))).T

    def maxpairs(self) -> int:
        return int(lib().orc_gcl_maxpairs(self.h))


FIELDS = {"x": (0, True), "v": (1, True), "rho": (2, False), "P": (3, False), "acc": (4, True), "sumW": (5, False),
          "sumgradW": (6, True), "rhoh": (7, False), "vh": (8, True), "xh": (9, True), "drho": (10, False),
          "cdx": (11, True), "dpc": (13, True)}


class SPHProblem:
    """SPHProblem(system, params, rho, ptype; v) on the CPU (src/sphproblem.jl:113-193)."""

    def __init__(self, system: CellList, params, rho, ptype, v=None):
        L = lib()
        self.system = system
        self.dim = system.dim
        self.n = system.n
        par = np.ascontiguousarray(params.as_array())
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        pt = np.ascontiguousarray(ptype, dtype=np.int32)
        vv = None
        if v is not None:
            vv = [np.ascontiguousarray(a, dtype=np.float64) for a in v]
        h = L.orc_sph_create(system.h, _dptr(par), _dptr(rho), pt.ctypes.data_as(C.POINTER(C.c_int32)),
                             _ptr_array(vv) if vv is not None else None)
        if not h:
            raise OracleError(L.orc_last_error().decode())
        self.h = h
        self.params = params

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sph_destroy(self.h)
            self.h = None

    def stepsolve(self, n: int = 1, timestepping: bool = False, timelims: Optional[Tuple[float, float]] = None):
        p = self.params
        if timelims is None:
            timelims = (np.sqrt(np.finfo(np.float64).eps), p.CFL * p.H / (3 * p.c0))
        upd = C.c_int64(0)
        mx = C.c_double(0)
        rc = lib().orc_sph_stepsolve(self.h, int(n), int(bool(timestepping)), float(timelims[0]),
                                     float(timelims[1]), C.byref(upd), C.byref(mx))
        if rc:
            raise OracleError(lib().orc_last_error().decode())
        return upd.value, mx.value

    def get(self, name: str) -> np.ndarray:
        code, vec = FIELDS[name]
        out = np.empty(self.n * (self.dim if vec else 1))
        lib().orc_sph_get_field(self.h, code, _dptr(out))
        return out.reshape(self.dim, self.n) if vec else out

    @property
    def dt(self) -> float:
        return lib().orc_sph_get_scalar(self.h, 0)

    @dt.setter
    def dt(self, v: float):
        lib().orc_sph_set_scalar(self.h, 0, float(v))

    @property
    def etime(self) -> float:
        return lib().orc_sph_get_scalar(self.h, 1)


def wendland_W(dim: int, u: float, hinv: float) -> float:
    return lib().orc_wendland_W(dim, u, hinv)


def wendland_dW(dim: int, u: float, hinv: float) -> float:
    return lib().orc_wendland_dW(dim, u, hinv)


def powfancy7th(x: float) -> float:
    return lib().orc_powfancy7th(x)


def dpckernlim(dim: int) -> float:
    return lib().orc_dpckernlim(dim)


def eos(rho: float, gamma: int, rho0: float, c0: float, ptype: int) -> float:
    return lib().orc_eos(rho, gamma, rho0, c0, ptype)


def omp_lib():
    """The multi-core CPU step (oracle/cpu_omp.cpp): the CPU baseline."""
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_PATH):
            build()
        L = C.CDLL(OMP_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.omp_sph_create.restype = P
        L.omp_sph_create.argtypes = [C.c_int, C.c_int64, C.POINTER(dp), dp, C.c_double, dp, dp,
                                     C.POINTER(C.c_int32), C.POINTER(dp)]
        L.omp_sph_destroy.argtypes = [P]
        L.omp_sph_stepsolve.argtypes = [P, C.c_int64, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int64), dp]
        L.omp_sph_get_field.argtypes = [P, C.c_int, dp]
        L.omp_sph_get_scalar.restype = C.c_double
        L.omp_sph_get_scalar.argtypes = [P, C.c_int]
        L.omp_sph_threads.argtypes = [P]
        L.omp_sph_pairs.restype = C.c_int64
        L.omp_sph_pairs.argtypes = [P]
        _omp = L
    return _omp


class OmpSPHProblem:
    """The SPH step on all host cores (OpenMP per-particle gather over a CSR
    neighbour list, FP64): same interface subset as SPHProblem.  The cell
    grid follows the reference's GPUCellList constructor over the points."""

    def __init__(self, points, cellsize, dist: float, params, rho, ptype, v=None):
        L = omp_lib()
        self.dim = len(points)
        pts = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
        self.n = pts[0].shape[0]
        cs = np.ascontiguousarray(np.broadcast_to(np.asarray(cellsize, dtype=np.float64), (self.dim,)))
        par = np.ascontiguousarray(params.as_array())
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        pt = np.ascontiguousarray(ptype, dtype=np.int32)
        vv = None if v is None else [np.ascontiguousarray(a, dtype=np.float64) for a in v]
        h = L.omp_sph_create(self.dim, self.n, _ptr_array(pts), _dptr(cs), float(dist), _dptr(par), _dptr(rho),
                             pt.ctypes.data_as(C.POINTER(C.c_int32)), _ptr_array(vv) if vv is not None else None)
        if not h:
            raise OracleError(L.orc_last_error().decode())
        self.h = h
        self.params = params

    def __del__(self):
        if getattr(self, "h", None):
            omp_lib().omp_sph_destroy(self.h)
            self.h = None

    @property
    def threads(self) -> int:
        return omp_lib().omp_sph_threads(self.h)

    @property
    def pairs(self) -> int:
        return omp_lib().omp_sph_pairs(self.h)

    def stepsolve(self, n: int = 1, timestepping: bool = False, timelims: Optional[Tuple[float, float]] = None):
        p = self.params
        if timelims is None:
            timelims = (np.sqrt(np.finfo(np.float64).eps), p.CFL * p.H / (3 * p.c0))
        upd = C.c_int64(0)
        mx = C.c_double(0)
        rc = omp_lib().omp_sph_stepsolve(self.h, int(n), int(bool(timestepping)), float(timelims[0]),
                                         float(timelims[1]), C.byref(upd), C.byref(mx))
        if rc:
            raise OracleError(omp_lib().orc_last_error().decode())
        return upd.value, mx.value

    def get(self, name: str) -> np.ndarray:
        code, vec = FIELDS[name]
        out = np.empty(self.n * (self.dim if vec else 1))
        if omp_lib().omp_sph_get_field(self.h, code, _dptr(out)):
            raise OracleError(omp_lib().orc_last_error().decode())
        return out.reshape(self.dim, self.n) if vec else out

    @property
    def dt(self) -> float:
        return omp_lib().omp_sph_get_scalar(self.h, 0)

    @property
    def etime(self) -> float:
        return omp_lib().omp_sph_get_scalar(self.h, 1)
