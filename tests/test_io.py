"""Output formats (src/auxillary.jl:18-70, src/writevtk.jl:1-43): DataFrame / CSV
columns and the VTK PolyData + ParaView collection files, written from a
stand-in problem object (the writers only call the getters), then read back."""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "gpucelllistsph.jl_amd"))
from gclsph import io as gio  # noqa: E402
from gclsph import api  # noqa: E402


class FakeProblem:
    """Holds fields the getters return; etime for writetime."""

    def __init__(self, dim, n, seed=3):
        rng = np.random.default_rng(seed)
        self.dim, self.n, self.etime = dim, n, 0.125
        self.f = {
            "x": tuple(rng.random(n) for _ in range(dim)),
            "v": tuple(rng.standard_normal(n) for _ in range(dim)),
            "a": tuple(rng.standard_normal(n) for _ in range(dim)),
            "rho": 1000 + rng.standard_normal(n),
            "p": rng.random(n) * 1e4,
            "pt": rng.integers(0, 3, n).astype(np.int32),
            "sw": rng.random(n),
            "sgw": tuple(rng.standard_normal(n) for _ in range(dim)),
            "dpc": tuple(rng.standard_normal(n) * 1e-3 for _ in range(dim)),
        }


@pytest.fixture(autouse=True)
def fake_getters(monkeypatch):
    for name, key in [("get_points", "x"), ("get_velocity", "v"), ("get_acceleration", "a"), ("get_density", "rho"),
                      ("get_pressure", "p"), ("get_ptype", "pt"), ("get_sumw", "sw"), ("get_sumgradw", "sgw"),
                      ("get_dpccorr", "dpc")]:
        monkeypatch.setattr(api, name, lambda prob, key=key: prob.f[key])


@pytest.mark.parametrize("dim", [2, 3])
def test_makedf_columns_follow_reference_order(dim):
    prob = FakeProblem(dim, 50)
    allv = ["Density", "Pressure", "Acceleration", "Velocity", "∑W", "∑∇W", "DPC", "Type"]
    df = gio.makedf(prob, allv)
    ax = ["x", "y", "z"][:dim]
    want = ax + ["Density", "Pressure"] + [f"Acceleration_{a}" for a in ax] + [f"Velocity_{a}" for a in ax] + \
        ["∑W"] + [f"∑∇W_{a}" for a in ax] + [f"DPC_{a}" for a in ax]
    assert list(df.columns) == want  # src/auxillary.jl:22-58 (Type is not a makedf column)
    np.testing.assert_array_equal(df["Velocity_y"].to_numpy(), prob.f["v"][1])
    np.testing.assert_array_equal(df["Density"].to_numpy(), prob.f["rho"])
    # default vtkvars: Density, Acceleration, Velocity
    assert list(gio.makedf(prob).columns) == ax + ["Density"] + [f"Acceleration_{a}" for a in ax] + \
        [f"Velocity_{a}" for a in ax]


def test_writecsv_roundtrip_and_wrong_path(tmp_path):
    import pandas as pd
    prob = FakeProblem(2, 40)
    f = tmp_path / "out.csv"
    gio.writecsv(prob, str(f), ["Density", "Velocity"])
    back = pd.read_csv(f, float_precision="round_trip")
    assert list(back.columns) == ["x", "y", "Density", "Velocity_x", "Velocity_y"]
    np.testing.assert_array_equal(back["x"].to_numpy(), prob.f["x"][0])  # %.17g round-trips FP64
    with pytest.raises(ValueError, match="Wrong path"):
        gio.writecsv(prob, str(tmp_path / "missing" / "out.csv"))


@pytest.mark.parametrize("dim", [2, 3])
def test_writevtk_polydata_roundtrip(tmp_path, dim):
    prob = FakeProblem(dim, 77)
    path = gio.writevtk(prob, str(tmp_path / "OUTPUT_00000"), ["Density", "Pressure", "Type", "Velocity", "DPC"])
    assert path.endswith(".vtp") and os.path.exists(path)
    root = ET.parse(path).getroot()
    assert root.get("type") == "PolyData" and root.get("compressor") == "vtkZLibDataCompressor"
    piece = root.find("PolyData/Piece")
    assert piece.get("NumberOfPoints") == "77" and piece.get("NumberOfVerts") == "0" and piece.get("NumberOfPolys") == "0"
    back = gio.read_vtp(path)
    pts = back["Points"]
    assert pts.shape == (77, 3)
    for k in range(dim):
        np.testing.assert_array_equal(pts[:, k], prob.f["x"][k])
    if dim == 2:
        assert not pts[:, 2].any()  # zero z in 2D
    np.testing.assert_array_equal(back["Density"], prob.f["rho"])
    np.testing.assert_array_equal(back["Type"], prob.f["pt"])
    assert back["Type"].dtype == np.int32
    assert back["Velocity"].shape == (77, dim)
    np.testing.assert_array_equal(back["Velocity"], np.stack(prob.f["v"], axis=1))
    np.testing.assert_array_equal(back["DPC"], np.stack(prob.f["dpc"], axis=1))
    assert "Acceleration" not in back


def test_paraview_collection_lists_timesteps(tmp_path):
    prob = FakeProblem(2, 10)
    pvd = gio.paraview_collection(str(tmp_path / "OUTPUT_PVC"))
    for i, t in enumerate([0.0, 0.5, 1.0]):
        prob.etime = t
        gio.writevtk(prob, str(tmp_path / f"OUTPUT_{i:05d}"), ["Density"], pvd=pvd, writetime=True)
    gio.writevtk(prob, str(tmp_path / "not_in_pvd"), ["Density"], pvd=pvd, writetime=False)
    path = gio.vtk_save(pvd)
    assert path.endswith("OUTPUT_PVC.pvd")
    ds = ET.parse(path).getroot().findall("Collection/DataSet")
    assert [float(d.get("timestep")) for d in ds] == [0.0, 0.5, 1.0]
    assert [d.get("file") for d in ds] == ["OUTPUT_00000.vtp", "OUTPUT_00001.vtp", "OUTPUT_00002.vtp"]


def test_vtk_binary_block_header():
    """Inline binary arrays: base64(UInt64 [1, n bytes, n bytes, compressed bytes]) + base64(zlib data)."""
    import base64
    import zlib
    a = np.arange(10, dtype=np.float64)
    txt = gio._encode(a)
    head = np.frombuffer(base64.b64decode(txt[:44]), dtype="<u8")
    assert list(head[:3]) == [1, 80, 80]
    data = base64.b64decode(txt[44:])
    assert len(data) == head[3]
    np.testing.assert_array_equal(np.frombuffer(zlib.decompress(data), dtype="<f8"), a)


def test_background_writer_order_and_content(tmp_path):
    """VtkWriter: files written off-thread, in submission order, registered in
    the collection in that order, with the values taken at submission time."""
    prob = FakeProblem(2, 300)
    pvd = gio.paraview_collection(str(tmp_path / "OUT"))
    w = gio.VtkWriter()
    expected = []
    for k in range(4):
        prob.etime = 0.1 * (k + 1)
        prob.f["rho"] = prob.f["rho"] + 1.0  # a new array each time, as the getters return
        expected.append(prob.f["rho"].copy())
        w.submit(prob, str(tmp_path / f"OUTPUT_{k:05d}"), ("Density", "Velocity"), pvd=pvd, writetime=True)
    paths = w.close()
    assert paths == [str(tmp_path / f"OUTPUT_{k:05d}.vtp") for k in range(4)]
    assert [t for t, _ in pvd.entries] == [0.1 * (k + 1) for k in range(4)]
    for k, p in enumerate(paths):
        d = gio.read_vtp(p)
        np.testing.assert_array_equal(d["Density"], expected[k])
        np.testing.assert_array_equal(d["Velocity"][:, 1], prob.f["v"][1])


def test_background_writer_reraises(tmp_path):
    prob = FakeProblem(2, 10)
    w = gio.VtkWriter()
    w.submit(prob, str(tmp_path / "missing_dir" / "OUTPUT_00000"), ("Density",))
    with pytest.raises(OSError):
        w.close()
