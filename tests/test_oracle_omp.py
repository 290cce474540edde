"""The multi-core CPU step (oracle/cpu_omp.cpp, the CPU baseline bench.py
times) against the serial oracle: the same per-pair arithmetic summed in a
different order, so FP64 results agree to rounding after many steps."""
import numpy as np
import pytest


def _both(O, sc, **over):
    for k, v in over.items():
        setattr(sc.params, k, v)
    oc = O.CellList(sc.points, sc.cellsize, sc.dist)
    ref = O.SPHProblem(oc, sc.params, sc.rho, sc.ptype, v=sc.v)
    omp = O.OmpSPHProblem(sc.points, sc.cellsize, sc.dist, sc.params, sc.rho, sc.ptype, v=sc.v)
    return ref, omp


def _close(ref, omp, dx, steps=None):
    assert np.abs(omp.get("x") - ref.get("x")).max() < 1e-11 * dx
    assert np.abs(omp.get("rho") / ref.get("rho") - 1).max() < 1e-12
    vs = np.abs(ref.get("v")).max()
    assert np.abs(omp.get("v") - ref.get("v")).max() < 1e-10 * max(vs, 1e-3)
    assert omp.etime == pytest.approx(ref.etime, rel=1e-13)
    assert omp.dt == pytest.approx(ref.dt, rel=1e-10)


def test_omp_step_matches_oracle_config0(O, S):
    sc = S.config0()
    ref, omp = _both(O, sc)
    for n in (6, 4):
        assert omp.stepsolve(n, True) == pytest.approx(ref.stepsolve(n, True), rel=1e-9)
    _close(ref, omp, sc.params.dx)
    assert omp.threads >= 1


def test_omp_step_matches_oracle_with_viscosity_and_surface_tension(O, S):
    sc = S.dam_break_2d(0.03, 0.6, 0.9, 2.0, 1.5, boundary_ptype=-1, dist_factor=1.2, name="uneven")
    ref, omp = _both(O, sc, nu=0.05, s=0.02, bound_l=1.5 * sc.params.dx, cspmn=0)
    for n in (5, 5):
        upd_r, mx_r = ref.stepsolve(n, True)
        upd_o, mx_o = omp.stepsolve(n, True)
        assert upd_o == upd_r and mx_o == pytest.approx(mx_r, rel=1e-9)
    _close(ref, omp, sc.params.dx)


def test_omp_step_3d(O, S):
    sc = S.config2(dx=0.05)
    ref, omp = _both(O, sc, dpc_l0=0.0, xsph_eps=0.0, cspmn=0)
    ref.stepsolve(4, True)
    omp.stepsolve(4, True)
    _close(ref, omp, sc.params.dx)


def test_omp_step_refuses_corrections(O, S):
    sc = S.config0()
    sc.params.xsph_eps = 0.5
    with pytest.raises(O.OracleError, match="DPC or XSPH"):
        O.OmpSPHProblem(sc.points, sc.cellsize, sc.dist, sc.params, sc.rho, sc.ptype)
    sc = S.config0()
    omp = O.OmpSPHProblem(sc.points, sc.cellsize, sc.dist, sc.params, sc.rho, sc.ptype)
    with pytest.raises(O.OracleError, match="CSPM"):
        omp.stepsolve(sc.params.cspmn)
