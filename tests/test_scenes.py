"""Scene generators: the RNG's published test vector and the fixture geometry."""
import os

import numpy as np
from conftest import GOLDEN


def test_splitmix64_test_vector(S):
    # SplitMix64 started from state 0: first output 0xE220A8397B1DCDAF
    assert int(S.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    u = S.uniform01(1, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.005
    z = S.normal(2, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def test_config0_matches_fixture_geometry(S):
    sc = S.config0()
    assert sc.info["fluid"] == 5000 and sc.info["boundary"] == 1491
    pts, rho, ptype = S.load_dualsphysics_csv(os.path.join(GOLDEN, "FluidPoints_Dp0.02.csv"),
                                              os.path.join(GOLDEN, "BoundaryPoints_Dp0.02.csv"))
    wall = ptype < 2
    fx = np.sort(np.round(pts[0][wall], 6) * 1000 + np.round(pts[1][wall], 6))
    nb = int(sc.info["boundary"])
    gx = np.sort(np.round(sc.points[0][-nb:], 6) * 1000 + np.round(sc.points[1][-nb:], 6))
    assert np.array_equal(fx, gx)  # the generated walls are the fixture's walls


def test_config3_size(S):
    sc = S.config3(dx=0.02)  # same construction at a coarse spacing
    assert sc.info["fluid"] == 200 * 200
    p = sc.params
    assert abs(p.H - 2 * 1.2 * np.sqrt(2) * 0.02) < 1e-15 and sc.dist == p.H
