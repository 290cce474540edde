"""Python mirror of GPUCellListSPH.jl's exported API (src/GPUCellListSPH.jl:9-17).

Julia's in-place `f!` functions are spelled `f_` here (update_, partialupdate_,
stepsolve_, timesolve_).  Arguments keep the reference's order and meaning;
vectors are tuples of per-component numpy arrays (the reference's
NTuple{D,CuArray} layout) in the caller's particle order.  Errors the reference
raises with `error(...)` raise ValueError here; soft conditions warn.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import warnings
import weakref
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, lib

_EPS = float(np.finfo(np.float64).eps)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ptrs(arrs):
    t = (C.POINTER(C.c_double) * len(arrs))()
    for k, a in enumerate(arrs):
        t[k] = _dp(a)
    return t


def _prec(precision) -> int:
    if precision in (64, "f64", "Float64", np.float64, float):
        return _lib.GCLSPH_F64
    if precision in (32, "f32", "Float32", np.float32):
        return _lib.GCLSPH_F32
    if precision in (33, "f32split", "fp32-split", "split"):
        return _lib.GCLSPH_F32SPLIT
    raise ValueError(f"unknown precision {precision!r}")


class WendlandC2:
    """SPHKernels.jl's WendlandC2(T, dim) — the kernel the reference uses."""

    def __init__(self, T=np.float64, dim: int = 2):
        if isinstance(T, int) and dim == 2 and T in (2, 3):
            T, dim = np.float64, T
        self.dim = int(dim)

    def __repr__(self):
        return f"WendlandC2({self.dim})"


# --------------------------------------------------------------------------
# GPUCellList (src/celllist.jl)
# --------------------------------------------------------------------------
class GPUCellList:
    """GPUCellList(points, cellsize, dist; MIN, MAX, mppcell, mpairs) — src/celllist.jl:22-88."""

    def __init__(self, points: Sequence[np.ndarray], cellsize, dist: float, MIN=None, MAX=None, mppcell: int = 0,
                 mpairs: int = 0, precision="f64"):
        lens = [len(p) for p in points]
        if any(l != lens[0] for l in lens):
            raise ValueError("length of points not equal")
        D = len(points)
        if D < 2 or D > 3:
            raise ValueError("wrong dimention")
        self._pts = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
        cs = np.ascontiguousarray(np.broadcast_to(np.asarray(cellsize, dtype=np.float64), (D,)))
        mn = None if MIN is None else np.ascontiguousarray(np.broadcast_to(np.asarray(MIN, np.float64), (D,)))
        mx = None if MAX is None else np.ascontiguousarray(np.broadcast_to(np.asarray(MAX, np.float64), (D,)))
        warn = C.c_int(0)
        h = C.c_void_p()
        check(lib().gcl_create(D, _prec(precision), lens[0], _ptrs(self._pts), _dp(cs), float(dist),
                               _dp(mn) if mn is not None else None, _dp(mx) if mx is not None else None,
                               int(mppcell), int(mpairs), C.byref(warn), C.byref(h)))
        self._h = h
        self.dim = D
        self.precision = _prec(precision)
        if warn.value & 1:
            warnings.warn("Cell size < dist, cell size set equal dist")
        if warn.value & 2:
            warnings.warn("List of pairs not full, call `update!(c::GPUCellList)` to extend pair list!")
        self._owner = None

    def __del__(self):
        h = getattr(self, "_h", None)
        ref = getattr(self, "_owner", None)
        owner = ref() if ref is not None else None
        if h and owner is None:
            lib().gcl_destroy(h)
            self._h = None

    # reference fields -----------------------------------------------------
    def _info(self) -> _lib.GCLInfo:
        inf = _lib.GCLInfo()
        check(lib().gcl_info(self._h, C.byref(inf)))
        return inf

    @property
    def n(self) -> int:
        return int(self._info().n)

    @property
    def dist(self) -> float:
        return float(self._info().dist)

    @property
    def cs(self) -> Tuple[float, ...]:
        return tuple(self._info().cellsize[: self.dim])

    @property
    def offset(self) -> Tuple[float, ...]:
        return tuple(self._info().offset[: self.dim])

    @property
    def grid(self) -> Tuple[int, ...]:
        return tuple(int(g) for g in self._info().grid[: self.dim])

    @property
    def mppcell(self) -> int:
        return int(self._info().mppcell)

    @property
    def pairsn(self) -> int:
        v = C.c_int64()
        check(lib().gcl_pairsn(self._h, C.byref(v)))
        return int(v.value)

    @property
    def capacity(self) -> int:
        v = C.c_int64()
        check(lib().gcl_capacity(self._h, C.byref(v)))
        return int(v.value)

    @property
    def points(self) -> Tuple[np.ndarray, ...]:
        n = self.n
        out = [np.empty(n) for _ in range(self.dim)]
        check(lib().gcl_get_points(self._h, _ptrs(out)))
        return tuple(out)

    @points.setter
    def points(self, pts):
        arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in pts]
        if len(arrs) != self.dim or any(len(a) != self.n for a in arrs):
            raise ValueError("points must be dim arrays of length n")
        check(lib().gcl_set_points(self._h, _ptrs(arrs)))

    @property
    def cellpnum(self) -> np.ndarray:
        g = self.grid
        out = np.zeros(int(np.prod(g)), dtype=np.int32)
        check(lib().gcl_get_cellpnum(self._h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out.reshape(tuple(reversed(g))).T  # column-major like the reference

    @property
    def pcell(self) -> np.ndarray:
        out = np.zeros(self.dim * self.n, dtype=np.int32)
        check(lib().gcl_get_pcell(self._h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out.reshape(self.dim, self.n)

    def maxpairs(self) -> int:
        v = C.c_int64()
        check(lib().gcl_maxpairs(self._h, C.byref(v)))
        return int(v.value)

    def __repr__(self):
        g = self.grid
        return (f"    Cell List (Dim: {self.dim})\n  Points number: {self.n} Type: "
                f"{'Float32' if self.precision == 32 else 'Float64'}\n  Pairs number: {self.pairsn}\n"
                f"  Cell grid: {'X'.join(str(x) for x in g)}\n  Cutoff: {self.dist}")


def update_(c: GPUCellList, fillzero: bool = True) -> None:
    """update!(c, fillzero) — src/celllist.jl:95-134."""
    check(lib().gcl_update(c._h, int(bool(fillzero))))


def partialupdate_(c: GPUCellList, fillzero: bool = True) -> None:
    """partialupdate!(c, fillzero) — src/celllist.jl:141-145."""
    check(lib().gcl_partialupdate(c._h, int(bool(fillzero))))


def neighborlist(c: GPUCellList) -> np.ndarray:
    """neighborlist(c) — the whole pair buffer (capacity x 2, int32, 1-based, zero padded)."""
    cap = c.capacity
    out = np.zeros((max(cap, 0), 2), dtype=np.int32)
    w = C.c_int64()
    if cap > 0:
        check(lib().gcl_get_pairs(c._h, out.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(w)))
    return out


NOALLOC_MAX_POINTS = 524288
NOALLOC_MAX_CELLS = 1048576


def neighborlist_noalloc(pts_dev: Sequence[int], n: int, cellsize, dist: float, MIN, MAX, pairs_dev: int,
                         capacity: int, stream: int = 0) -> int:
    """The pair list of GPUCellList(points, cellsize, dist; MIN, MAX) +
    neighborlistview without any device allocation (gcl_neighborlist_noalloc;
    no counterpart in the reference).  pts_dev: D device addresses of n
    Float64 coordinates; pairs_dev: device address of 2*capacity int32, which
    receives the first min(pairsn, capacity) pairs (1-based, i < j) in the
    order gcl_get_pairs gives.  Returns pairsn.  n <= NOALLOC_MAX_POINTS and
    at most NOALLOC_MAX_CELLS cells; calls on one device share a fixed
    scratch array."""
    D = len(pts_dev)
    ptrs = (C.c_void_p * D)(*[C.c_void_p(int(p)) for p in pts_dev])
    cs = np.ascontiguousarray(np.broadcast_to(np.asarray(cellsize, dtype=np.float64), (D,)))
    mn = np.ascontiguousarray(np.broadcast_to(np.asarray(MIN, dtype=np.float64), (D,)))
    mx = np.ascontiguousarray(np.broadcast_to(np.asarray(MAX, dtype=np.float64), (D,)))
    out = C.c_int64(0)
    check(lib().gcl_neighborlist_noalloc(D, int(n), ptrs, _dp(cs), float(dist), _dp(mn), _dp(mx),
                                         C.c_void_p(int(pairs_dev)), int(capacity), C.byref(out),
                                         C.c_void_p(int(stream)) if stream else None))
    return int(out.value)


def neighborlistview(c: GPUCellList) -> np.ndarray:
    """neighborlistview(c) — view(c.pairs, 1:c.pairsn); BoundsError if pairsn > capacity."""
    n = c.pairsn
    cap = c.capacity
    if n > cap:
        raise IndexError(f"attempt to access {cap}-element pair list at index [1:{n}]")
    return neighborlist(c)[:n]


# --------------------------------------------------------------------------
# SPHProblem (src/sphproblem.jl)
# --------------------------------------------------------------------------
class SPHProblem:
    """SPHProblem(system, dx, h, H, sphkernel, ρ, ptype, ρ₀, m₀, Δt, α, c₀, γ, δᵩ, CFL; v, g, s)
    — src/sphproblem.jl:113-193.  `precision` selects the Float64 path ("f64") or
    the FP32-split path ("f32split")."""

    _MUTABLE = _lib.PARAMS

    def __init__(self, system: GPUCellList, dx, h, H, sphkernel, rho, ptype, rho0, m0, dt, alpha, c0, gamma, delta,
                 CFL, v=None, g=9.81, s=0.0, precision="f64"):
        N = system.n
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        pt = np.ascontiguousarray(ptype, dtype=np.int32)
        if len(rho) != N:
            raise ValueError(f"length(ρ) ({len(rho)}) != N ({N})")
        if len(pt) != N:
            raise ValueError(f"length(ptype) ({len(pt)}) != N ({N})")
        vv = None
        if v is not None:
            vv = [np.ascontiguousarray(a, dtype=np.float64) for a in v]
            if len(vv) != system.dim or any(len(a) != N for a in vv):
                raise ValueError(f"length(v) ({[len(a) for a in vv]}) != N ({N})")
        if sphkernel is not None and not isinstance(sphkernel, WendlandC2):
            # the kernels evaluate SPHKernels.jl's WendlandC2 only (the one the reference's examples use)
            raise NotImplementedError(f"SPH kernel {sphkernel!r}: only WendlandC2 is implemented")
        if sphkernel is not None and sphkernel.dim != system.dim:
            raise ValueError("SPH kernel dimension differs from the system's")
        par = {"dx": dx, "h": h, "H": H, "rho0": rho0, "m0": m0, "dt": dt, "alpha": alpha, "c0": c0, "gamma": gamma,
               "delta": delta, "CFL": CFL, "g": g, "s": s, "nu": 0.0, "dpc_l0": 0.0, "dpc_pmin": 1.0,
               "dpc_pmax": 10000.0, "dpc_lambda": 0.01, "xsph_eps": 0.0, "cspmn": 15, "bound_D": 0.4,
               "bound_l": 2 * dx, "nui": system.dist - H, "etime": 0.0}
        arr = np.array([float(par[k]) for k in _lib.PARAMS], dtype=np.float64)
        hh = C.c_void_p()
        check(lib().sph_create(system._h, _prec(precision), _dp(arr), _dp(rho),
                               pt.ctypes.data_as(C.POINTER(C.c_int32)), _ptrs(vv) if vv is not None else None,
                               C.byref(hh)))
        object.__setattr__(self, "_h", hh)
        object.__setattr__(self, "system", system)
        object.__setattr__(self, "dim", system.dim)
        object.__setattr__(self, "sphkernel", sphkernel)
        object.__setattr__(self, "precision", _prec(precision))
        system._owner = weakref.ref(self)

    @classmethod
    def from_params(cls, system: GPUCellList, params, rho, ptype, v=None, precision="f64") -> "SPHProblem":
        """Builds a problem from a scenes.SPHParams record (all mutable fields set)."""
        p = params
        prob = cls(system, p.dx, p.h, p.H, WendlandC2(np.float64, system.dim), rho, ptype, p.rho0, p.m0, p.dt,
                   p.alpha, p.c0, p.gamma, p.delta, p.CFL, v=v, g=p.g, s=p.s, precision=precision)
        for k in ("nu", "dpc_l0", "dpc_pmin", "dpc_pmax", "dpc_lambda", "xsph_eps", "cspmn", "bound_D", "bound_l",
                  "nui", "etime"):
            setattr(prob, k, getattr(p, k))
        return prob

    def __getattr__(self, name):
        if name in _lib.PARAM_INDEX:
            v = C.c_double()
            check(lib().sph_get_param(self._h, _lib.PARAM_INDEX[name], C.byref(v)))
            return int(v.value) if name in ("gamma", "cspmn") else v.value
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _lib.PARAM_INDEX:
            check(lib().sph_set_param(self._h, _lib.PARAM_INDEX[name], float(value)))
        else:
            object.__setattr__(self, name, value)

    def __del__(self):
        h = self.__dict__.get("_h")
        if h:
            lib().sph_destroy(h)
            self.__dict__["_h"] = None
            sysobj = self.__dict__.get("system")
            if sysobj is not None:
                sysobj._owner = None

    def field(self, code: int, vector: bool) -> np.ndarray:
        n = self.system.n
        out = np.empty(n * (self.dim if vector else 1))
        check(lib().sph_get_field(self._h, code, _dp(out)))
        return tuple(out.reshape(self.dim, n)) if vector else out

    def set_timing(self, on: bool = True):
        check(lib().sph_set_timing(self._h, int(bool(on))))

    def timing(self):
        ms = (C.c_double * 8)()
        cnt = (C.c_int64 * 8)()
        ns = C.c_int()
        check(lib().sph_get_timing(self._h, ms, cnt, 8, C.byref(ns)))
        names = ("rebuild", "pass1", "pass2", "corrections", "finalize")
        return {names[k]: (ms[k], int(cnt[k])) for k in range(ns.value)}

    def pass_variants(self) -> dict:
        """The pair-pass variants the autotuner chose for this problem's own
        calls (variant = codegen + 4·blocksize index + 12·occupancy cap) and
        their block sizes; -1 until tuned."""
        out = (C.c_int32 * 4)()
        check(lib().sph_get_pass_variants(self._h, out))
        return {"pass1": int(out[0]), "pass2": int(out[1]), "pass1_threads": int(out[2]),
                "pass2_threads": int(out[3])}

    def __repr__(self):
        s = self.s
        dpc = f"l₀ = {self.dpc_l0} , λ = {self.dpc_lambda}" if self.dpc_l0 > 0 and self.dpc_lambda > 0 else "not used"
        return (f"  SPH Problem \n{self.system!r}\n  h: {self.h}\n  H: {self.H}\n  SPH Kernel: {self.sphkernel}\n"
                f"  E Time: {self.etime}\n  Surface tension: {s if s > 0 else 'not used'}\n  DPC: {dpc}")


def default_timelims(prob: SPHProblem) -> Tuple[float, float]:
    """(sqrt(eps()), CFL*H/3c₀) — src/sphproblem.jl:207."""
    return (math.sqrt(_EPS), prob.CFL * prob.H / (3 * prob.c0))


def stepsolve_(prob: SPHProblem, n: int = 1, timestepping: bool = False,
               timelims: Optional[Tuple[float, float]] = None) -> Tuple[int, float]:
    """stepsolve!(prob, n; timestepping, timelims) — src/sphproblem.jl:204-345.  Returns (updaten, maxcΔxout)."""
    if timelims is None:
        timelims = default_timelims(prob)
    upd = C.c_int64()
    mx = C.c_double()
    check(lib().sph_stepsolve(prob._h, int(n), int(bool(timestepping)), float(timelims[0]), float(timelims[1]),
                              C.byref(upd), C.byref(mx)))
    return int(upd.value), float(mx.value)


class SPHBatch:
    """Independent problems of one dimension and precision stepped together:
    every stage of the step is one launch over all of them (the batched
    ensemble).  Each problem keeps its own Δt, rebuilds, time and corrections
    and ends as if stepped alone; read results through the problems.  While a
    batch exists its problems cannot be destroyed (the batch keeps them alive)."""

    def __init__(self, problems: Sequence[SPHProblem]):
        probs = list(problems)
        arr = (C.c_void_p * len(probs))(*[p._h for p in probs])
        hh = C.c_void_p()
        check(lib().sph_batch_create(arr, len(probs), C.byref(hh)))
        self._h = hh
        self.problems = probs

    def __len__(self):
        return len(self.problems)

    def pass_variants(self) -> dict:
        """The pair-pass variants the autotuner chose for this batch (see
        SPHProblem.pass_variants)."""
        out = (C.c_int32 * 4)()
        check(lib().sph_batch_get_pass_variants(self._h, out))
        return {"pass1": int(out[0]), "pass2": int(out[1]), "pass1_threads": int(out[2]),
                "pass2_threads": int(out[3])}

    def __del__(self):
        h = self.__dict__.get("_h")
        if h:
            lib().sph_batch_destroy(h)
            self._h = None


def batch_stepsolve_(batch: SPHBatch, n: int = 1, timestepping: bool = False,
                     timelims: Optional[Sequence[Tuple[float, float]]] = None) -> Tuple[np.ndarray, np.ndarray]:
    """stepsolve!(prob, n; timestepping, timelims) for every problem of the batch
    in one launch per stage.  timelims: one (min, max) pair per problem, or None
    for each problem's default.  Returns (updaten, maxcΔxout) per problem."""
    k = len(batch)
    upd = np.zeros(k, dtype=np.int64)
    mx = np.zeros(k, dtype=np.float64)
    t0 = t1 = None
    if timelims is not None:
        t0 = np.ascontiguousarray([float(t[0]) for t in timelims], dtype=np.float64)
        t1 = np.ascontiguousarray([float(t[1]) for t in timelims], dtype=np.float64)
        if len(t0) != k:
            raise ValueError("one timelims pair per problem")
    check(lib().sph_batch_stepsolve(batch._h, int(n), int(bool(timestepping)), None if t0 is None else _dp(t0),
                                    None if t1 is None else _dp(t1), upd.ctypes.data_as(C.POINTER(C.c_int64)),
                                    _dp(mx)))
    return upd, mx


def timesolve_(prob: SPHProblem, batch: int = 10, timeframe: float = 1.0, writetime: float = 0, path=None,
               pvc: bool = False, vtkvars=("Acceleration", "Velocity", "Pressure"), timestepping: bool = False,
               timelims: Optional[Tuple[float, float]] = None, anim: bool = False, verbose: bool = True):
    """timesolve!(prob; batch, timeframe, writetime, path, pvc, vtkvars, …) —
    src/sphproblem.jl:420-511.  Steps in batches of `batch` until the simulated
    time passes `timeframe`.  With writetime > 0 and a path, writes
    OUTPUT_<nnnnn>.vtp files (nnnnn = the batch counter, as the reference names
    them: the first write after the start reuses OUTPUT_00000) whenever the
    time passes the next multiple of writetime, and with pvc a ParaView
    collection OUTPUT_PVC.pvd.  Each file is compressed and written on a
    background thread while the next batch steps (the fields are copied to
    the host first); all files are complete when the call returns.  The
    animation branch (Plots.jl) is not part of this library: anim=True raises
    NotImplementedError."""
    from .io import VtkWriter, paraview_collection, vtk_save
    if anim:
        raise NotImplementedError("animation output is not implemented; write VTK files (path, pvc) instead")
    lim = prob.CFL * prob.H / (3 * prob.c0)
    if timelims is None:
        timelims = (math.sqrt(_EPS), lim)
    if timelims[1] > lim:
        warnings.warn(f"Maximum dt limit ({timelims[1]}) > CFL*H/3c₀ ({lim})")
    if timestepping:
        timelims = (max(timelims[0], _EPS), min(timelims[1], lim))
    if verbose:
        print("    Start simulation...")
        print(f"Timestepping: {timestepping}")
        if timestepping:
            print(f"Δt limitss: {timelims}")
        print(f"Batch: {batch}")
        print(f"NUI: {prob.nui}")
    nt = prob.etime + writetime
    i = 0
    pvd = None
    write = writetime > 0 and path is not None
    writer = VtkWriter() if write else None
    try:
        if write:
            if pvc:
                pvd = paraview_collection(os.path.join(path, "OUTPUT_PVC"))
            writer.submit(prob, os.path.join(path, f"OUTPUT_{i:05d}"), vtkvars, pvd=pvd, writetime=True)
        diag = None
        prog = _Progress("Calculating...:") if verbose else None
        while prob.etime <= timeframe:
            diag = stepsolve_(prob, batch, timestepping=timestepping, timelims=timelims)
            if writetime > 0 and nt < prob.etime:
                nt += writetime
                if path is not None:
                    writer.submit(prob, os.path.join(path, f"OUTPUT_{i:05d}"), vtkvars, pvd=pvd, writetime=True)
            i += 1
            if prog is not None:
                prog.next(iter=i, time=prob.etime, Δt=prob.dt, updn=diag[0], dxpncu=diag[1])
    finally:
        if writer is not None:
            writer.close()
    if write and pvc:
        vtk_save(pvd)
    if prog is not None:
        prog.finish()
    return diag


class _Progress:
    """Per-batch progress of timesolve_, the values ProgressMeter shows in the
    reference (src/sphproblem.jl:456,498,510: `ProgressUnknown` with
    showvalues iter, time, Δt, updn, dxpncu, then `finish!`), one line per
    batch on stderr with the batch rate."""

    def __init__(self, desc: str, stream=None):
        import sys
        import time
        self.desc, self.stream, self.count = desc, stream or sys.stderr, 0
        self.t0 = time.perf_counter()

    def _rate(self):
        import time
        el = time.perf_counter() - self.t0
        return el, (self.count / el if el > 0 else 0.0)

    def next(self, **values):
        self.count += 1
        el, rate = self._rate()
        vals = "  ".join(f"{k}: {v}" for k, v in values.items())
        print(f"{self.desc} {self.count} \t Time: {el:.2f} s ({rate:.2f} it/s)  {vals}", file=self.stream,
              flush=True)

    def finish(self):
        el, rate = self._rate()
        print(f"{self.desc} {self.count} \t Time: {el:.2f} s ({rate:.2f} it/s) done", file=self.stream, flush=True)


def get_points(prob: SPHProblem):
    return prob.field(_lib.FIELD_X, True)


def get_velocity(prob: SPHProblem):
    return prob.field(_lib.FIELD_V, True)


def get_density(prob: SPHProblem):
    return prob.field(_lib.FIELD_RHO, False)


def get_pressure(prob: SPHProblem):
    return prob.field(_lib.FIELD_P, False)


def get_acceleration(prob: SPHProblem):
    return prob.field(_lib.FIELD_ACC, True)


def get_sumw(prob: SPHProblem):
    return prob.field(_lib.FIELD_SUMW, False)


def get_sumgradw(prob: SPHProblem):
    return prob.field(_lib.FIELD_SUMGRADW, True)


def get_ptype(prob: SPHProblem):
    return prob.field(_lib.FIELD_PTYPE, False).astype(np.int32)


def get_dpccorr(prob: SPHProblem):
    """get_dpccorr (src/sphproblem.jl:368-370).  The reference's getter reads a
    field its SPHProblem does not define; this returns what the name says: the
    per-particle velocity correction Δv_dpc of the last step's dpcreg!
    (src/gpukernels_sph.jl:1500-1553), zero when the last call ran without DPC."""
    return prob.field(_lib.FIELD_DPC, True)


def get_simtime(prob: SPHProblem) -> float:
    return prob.etime


def get_dt(prob: SPHProblem) -> float:
    return prob.dt


def get_drho(prob: SPHProblem):
    """prob.∑∂ρ∂t of the last step."""
    return prob.field(_lib.FIELD_DRHO, False)
