"""Caller-visible outputs on the GPU path: get_dpccorr against the oracle's
last dpcreg! correction, and the VTK / CSV writers and timesolve!'s output
schedule on a real problem (files read back and compared with the getters)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest
from cases import golden_cases

pytestmark = pytest.mark.gpu


def _pair(G, O, sc, precision="f64"):
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    oc = O.CellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision=precision)
    ref = O.SPHProblem(oc, sc.params, sc.rho, sc.ptype, v=sc.v)
    return cl, oc, prob, ref


@pytest.mark.parametrize("case", ["dam2d_allphys", "dam3d_allphys"])
def test_dpccorr_matches_oracle(G, O, case):
    sc = golden_cases()[case][0]
    cl, oc, prob, ref = _pair(G, O, sc)
    G.stepsolve_(prob, 3, timestepping=True)
    ref.stepsolve(3, True)
    got = np.array(G.get_dpccorr(prob))
    want = ref.get("dpc")
    assert got.shape == want.shape
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(got - want).max() < 1e-9 * scale
    # a user update! reorders the state; the correction follows its particles
    G.update_(cl)
    np.testing.assert_array_equal(np.array(G.get_dpccorr(prob)), got)


def test_dpccorr_zero_without_dpc(G, S):
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype)
    G.stepsolve_(prob, 2)
    assert not np.array(G.get_dpccorr(prob)).any()


@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_writers_match_getters(G, S, tmp_path, precision):
    sc = golden_cases()["dam2d_allphys"][0]
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision=precision)
    G.stepsolve_(prob, 4, timestepping=True)
    path = G.writevtk(prob, str(tmp_path / "state"), ["Density", "Pressure", "Type", "Acceleration", "Velocity", "∑W",
                                                      "∑∇W", "DPC"])
    back = G.read_vtp(path)
    np.testing.assert_array_equal(back["Points"][:, :2], np.stack(G.get_points(prob), axis=1))
    np.testing.assert_array_equal(back["Density"], G.get_density(prob))
    np.testing.assert_array_equal(back["Pressure"], G.get_pressure(prob))
    np.testing.assert_array_equal(back["Type"], G.get_ptype(prob))
    np.testing.assert_array_equal(back["Velocity"], np.stack(G.get_velocity(prob), axis=1))
    np.testing.assert_array_equal(back["Acceleration"], np.stack(G.get_acceleration(prob), axis=1))
    np.testing.assert_array_equal(back["∑W"], G.get_sumw(prob))
    np.testing.assert_array_equal(back["∑∇W"], np.stack(G.get_sumgradw(prob), axis=1))
    np.testing.assert_array_equal(back["DPC"], np.stack(G.get_dpccorr(prob), axis=1))
    df = G.makedf(prob, ["Density", "Velocity"])
    np.testing.assert_array_equal(df["Velocity_x"].to_numpy(), G.get_velocity(prob)[0])
    G.writecsv(prob, str(tmp_path / "state.csv"))
    assert (tmp_path / "state.csv").read_text().splitlines()[0] == \
        "x,y,Density,Acceleration_x,Acceleration_y,Velocity_x,Velocity_y"


def test_timesolve_writes_vtk_schedule(G, S, tmp_path):
    """timesolve!(…; writetime, path, pvc): OUTPUT_<batch counter>.vtp files and the
    OUTPUT_PVC.pvd collection (src/sphproblem.jl:444-490)."""
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype)
    dt = prob.dt
    G.timesolve_(prob, batch=5, timeframe=30 * dt, writetime=10 * dt, path=str(tmp_path), pvc=True,
                 vtkvars=["Density", "Velocity"], verbose=False)
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".vtp"))
    ds = ET.parse(tmp_path / "OUTPUT_PVC.pvd").getroot().findall("Collection/DataSet")
    times = [float(d.get("timestep")) for d in ds]
    # an initial write at t=0, then one per passed multiple of writetime, named by the batch counter
    assert times[0] == 0.0 and times == sorted(times) and len(times) >= 3
    assert set(d.get("file") for d in ds) == set(files)
    assert files[0] == "OUTPUT_00000.vtp"
    last = G.read_vtp(str(tmp_path / ds[-1].get("file")))
    np.testing.assert_array_equal(last["Density"], G.get_density(prob))
    with pytest.raises(NotImplementedError):
        G.timesolve_(prob, batch=1, timeframe=prob.etime, anim=True, verbose=False)
