"""Loader for libgclsph.so (the C ABI declared in include/gclsph.h).

The library is built in-tree (gpucelllistsph.jl_amd/lib/libgclsph.so).  There
is no CPU fallback: if the library is missing, or no GPU is present, calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# GCLSPH_LIB names another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GCLSPH_LIB") or os.path.join(PKG_ROOT, "lib", "libgclsph.so")

GCLSPH_F64 = 64
GCLSPH_F32 = 32
GCLSPH_F32SPLIT = 33

ERRORS = {-1: ValueError, -2: RuntimeError, -3: RuntimeError, -4: RuntimeError, -5: NotImplementedError}

PARAMS = ("dx", "h", "H", "rho0", "m0", "dt", "alpha", "c0", "gamma", "delta", "CFL", "g", "s", "nu", "dpc_l0",
          "dpc_pmin", "dpc_pmax", "dpc_lambda", "xsph_eps", "cspmn", "bound_D", "bound_l", "nui", "etime")
PARAM_INDEX = {k: i for i, k in enumerate(PARAMS)}

FIELD_X, FIELD_V, FIELD_RHO, FIELD_P, FIELD_ACC, FIELD_SUMW, FIELD_SUMGRADW = 0, 1, 2, 3, 4, 5, 6
FIELD_DRHO, FIELD_PTYPE, FIELD_DPC = 10, 12, 13


class GCLInfo(C.Structure):
    _fields_ = [("dim", C.c_int32), ("prec", C.c_int32), ("n", C.c_int64), ("grid", C.c_int64 * 3),
                ("offset", C.c_double * 3), ("cellsize", C.c_double * 3), ("dist", C.c_double),
                ("mppcell", C.c_int64)]


_lib = None


def build(force: bool = False, jobs: int = 4) -> str:
    """Compiles libgclsph.so for gfx950 with hipcc (works without a GPU)."""
    cmd = ["make", "-C", PKG_ROOT, f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libgclsph.so not built ({LIB_PATH}); run gclsph.build() or `make -C {PKG_ROOT}`")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ddp = C.POINTER(dp)
    i64p = C.POINTER(C.c_int64)
    L.gclsph_last_error.restype = C.c_char_p
    L.gclsph_version.restype = C.c_char_p
    sigs = {
        "gcl_create": [C.c_int, C.c_int, C.c_int64, ddp, dp, C.c_double, dp, dp, C.c_int64, C.c_int64,
                       C.POINTER(C.c_int), C.POINTER(P)],
        "gcl_destroy": [P],
        "gcl_update": [P, C.c_int],
        "gcl_partialupdate": [P, C.c_int],
        "gcl_set_points": [P, ddp],
        "gcl_get_points": [P, ddp],
        "gcl_pairsn": [P, i64p],
        "gcl_capacity": [P, i64p],
        "gcl_get_pairs": [P, C.POINTER(C.c_int32), C.c_int64, i64p],
        "gcl_info": [P, C.POINTER(GCLInfo)],
        "gcl_get_cellpnum": [P, C.POINTER(C.c_int32)],
        "gcl_get_pcell": [P, C.POINTER(C.c_int32)],
        "gcl_maxpairs": [P, i64p],
        "sph_create": [P, C.c_int, dp, dp, C.POINTER(C.c_int32), ddp, C.POINTER(P)],
        "sph_destroy": [P],
        "sph_set_param": [P, C.c_int, C.c_double],
        "sph_get_param": [P, C.c_int, dp],
        "sph_stepsolve": [P, C.c_int64, C.c_int, C.c_double, C.c_double, i64p, dp],
        "sph_get_field": [P, C.c_int, dp],
        "sph_set_timing": [P, C.c_int],
        "sph_get_timing": [P, dp, i64p, C.c_int, C.POINTER(C.c_int)],
        "sph_batch_create": [C.POINTER(P), C.c_int32, C.POINTER(P)],
        "sph_batch_destroy": [P],
        "sph_batch_stepsolve": [P, C.c_int64, C.c_int, dp, dp, i64p, dp],
        "sph_get_pass_variants": [P, C.POINTER(C.c_int32)],
        "sph_batch_get_pass_variants": [P, C.POINTER(C.c_int32)],
        "gcl_neighborlist_noalloc": [C.c_int, C.c_int64, C.POINTER(C.c_void_p), dp, C.c_double, dp, dp, C.c_void_p,
                                     C.c_int64, i64p, C.c_void_p],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        msg = lib().gclsph_last_error().decode()
        raise ERRORS.get(rc, RuntimeError)(msg)


EXPORTED = ("gclsph_last_error", "gclsph_version", "gcl_create", "gcl_destroy", "gcl_update", "gcl_partialupdate",
            "gcl_set_points", "gcl_get_points", "gcl_pairsn", "gcl_capacity", "gcl_get_pairs", "gcl_info",
            "gcl_get_cellpnum", "gcl_get_pcell", "gcl_maxpairs", "sph_create", "sph_destroy", "sph_set_param",
            "sph_get_param", "sph_stepsolve", "sph_get_field", "sph_set_timing", "sph_get_timing", "sph_batch_create",
            "sph_batch_destroy", "sph_batch_stepsolve", "gcl_neighborlist_noalloc", "sph_get_pass_variants",
            "sph_batch_get_pass_variants")
