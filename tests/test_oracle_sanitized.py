"""The CPU oracle and the OpenMP CPU baseline under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C oracle sanitize` builds both with
-fsanitize=address,undefined (UB fatal), and a Python process started with the
ASan runtime preloaded drives them through the cell list (update!,
partialupdate!, мaxpairs, a NaN coordinate, points outside MIN/MAX), SPH steps
with every physics term and correction in 2D and 3D, and the OpenMP step.  Any
sanitizer report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
import numpy as np
import oracle as O
from gclsph import scenes as S

# cell list: fixture-sized random points, update!, partialupdate!, мaxpairs
pts = S.uniform_points(3, 4000, 2)
cl = O.CellList(pts, 0.03, 0.03)
cl.update()
cl.set_points(S.jitter(pts, 3, 1, sigma=0.002))
cl.partialupdate()
assert cl.info()["pairsn"] > 0
cl.maxpairs()
# a NaN coordinate and points outside the grid bounds
bad = (np.concatenate([pts[0][:500], [np.nan, 5.0, -3.0]]), np.concatenate([pts[1][:500], [0.5, 0.5, 0.5]]))
O.CellList(bad, 0.05, 0.05).update()
# 3D cell list
p3 = S.uniform_points(4, 3000, 3)
O.CellList(p3, 0.08, 0.08).update()

# SPH steps: 2D with viscosity, surface tension, DPC, XSPH and CSPM; 3D likewise
sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5, boundary_ptype=-1, dist_factor=1.1)
sc.params.nu = 0.1
sc.params.s = 0.05
sc.params.dpc_l0 = 1.1 * sc.params.dx
sc.params.dpc_lambda = 0.005
sc.params.dpc_pmax = 36000.0
sc.params.xsph_eps = 0.5
sc.params.cspmn = 2
c2 = O.CellList(sc.points, sc.cellsize, sc.dist)
p2 = O.SPHProblem(c2, sc.params, sc.rho, sc.ptype)
p2.stepsolve(6, True)
for f in ("x", "v", "rho", "P", "acc", "sumW", "sumgradW", "drho", "dpc"):
    assert np.all(np.isfinite(np.asarray(p2.get(f))))
sc3 = S.config2(dx=0.1)
c3 = O.CellList(sc3.points, sc3.cellsize, sc3.dist)
p3s = O.SPHProblem(c3, sc3.params, sc3.rho, sc3.ptype)
p3s.stepsolve(3, True)

# the OpenMP step (2 threads)
sc0 = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5, boundary_ptype=-1)
om = O.OmpSPHProblem(sc0.points, sc0.cellsize, sc0.dist, sc0.params, sc0.rho, sc0.ptype, None)
om.stepsolve(3, True)
print("sanitized oracle ok")
"""


def _asan_runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_address_and_undefined_sanitizers():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    build = os.path.join(ROOT, "oracle", "_build")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "GCLSPH_ORACLE_LIB": os.path.join(build, "liboracle_san.so"),
        "GCLSPH_ORACLE_OMP_LIB": os.path.join(build, "libcpuomp_san.so"),
        "OMP_NUM_THREADS": "2",
        "PYTHONPATH": os.pathsep.join([os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gpucelllistsph.jl_amd")]),
    })
    r = subprocess.run([sys.executable, "-c", DRIVER], env=env, capture_output=True, text=True, timeout=600)
    report = r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, report
    assert r.returncode == 0, report
    assert "sanitized oracle ok" in r.stdout
