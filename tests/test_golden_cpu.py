"""The oracle reproduces its committed golden vectors bit for bit."""
import os

import numpy as np
import pytest
from cases import golden_cases
from conftest import GOLDEN

FIELDS = ("x", "v", "rho", "P", "acc", "drho", "sumW", "sumgradW")


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_oracle_matches_golden(O, name):
    sc, steps, ts = golden_cases()[name]
    ref = np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"))
    cl = O.CellList(sc.points, sc.cellsize, sc.dist)
    assert np.array_equal(np.unique(cl.neighborlistview(), axis=0), ref["pairs0"])
    sph = O.SPHProblem(cl, sc.params, sc.rho, sc.ptype, v=sc.v)
    upd, mx = sph.stepsolve(steps, ts)
    assert upd == ref["updaten"] and mx == ref["maxcdx"]
    assert sph.dt == ref["dt"] and sph.etime == ref["etime"]
    for f in FIELDS:
        assert np.array_equal(sph.get(f), ref[f]), f
    assert np.array_equal(np.unique(cl.neighborlistview(), axis=0), ref["pairs_end"])
