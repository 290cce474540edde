"""Small seeded scenes shared by the golden-vector generator and the tests."""
import numpy as np

from gclsph import scenes as S


def golden_cases():
    """name -> (scene, steps, timestepping); shared with the tests."""
    cases = {}
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5, boundary_ptype=-1, name="small-2d")
    cases["dam2d"] = (sc, 4, True)
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5, dist_factor=1.1, name="small-2d-allphys")
    p = sc.params
    # dpc_l0 = 1.1 dx so that DPC acts on nearest neighbours (its kernel vanishes beyond l0)
    p.nu, p.s, p.dpc_l0, p.dpc_lambda, p.dpc_pmax, p.xsph_eps, p.cspmn = 0.2, 0.05, 0.055, 0.01, 36000.0, 0.5, 2
    cases["dam2d_allphys"] = (sc, 5, True)
    sc = S.dam_break_3d(0.05, (0.3, 0.3, 0.2), (0.8, 0.5, 0.3), name="small-3d")
    p = sc.params
    p.nu, p.s, p.dpc_l0, p.dpc_lambda, p.dpc_pmax, p.xsph_eps, p.cspmn = 0.2, 0.05, 0.055, 0.005, 36000.0, 0.5, 2
    cases["dam3d_allphys"] = (sc, 3, True)
    return cases


