"""gclsph — MI355X (gfx950) cell list + weakly-compressible SPH step.

Python mirror of GPUCellListSPH.jl's API over libgclsph.so.  Import with
`gpucelllistsph.jl_amd` on sys.path (the directory name is not a Python
identifier).
"""
from ._lib import GCLSPH_F32, GCLSPH_F32SPLIT, GCLSPH_F64, LIB_PATH, build, lib
from .api import (GPUCellList, SPHBatch, SPHProblem, WendlandC2, batch_stepsolve_, default_timelims,
                  get_acceleration, get_density, get_dpccorr, get_drho, get_dt, get_points, get_pressure, get_ptype,
                  get_simtime, get_sumgradw, get_sumw, get_velocity, neighborlist, neighborlist_noalloc,
                  neighborlistview, partialupdate_,
                  stepsolve_, timesolve_, update_)
from .io import (VtkWriter, collect_vtk, create_vtp_file, makedf, paraview_collection, read_vtp, vtk_save,
                 writecsv, writevtk)

__all__ = ["GPUCellList", "SPHProblem", "SPHBatch", "batch_stepsolve_", "WendlandC2", "update_", "partialupdate_",
           "neighborlist", "neighborlist_noalloc", "neighborlistview", "stepsolve_", "timesolve_", "get_points", "get_velocity",
           "get_density", "get_pressure", "get_acceleration", "get_sumw", "get_sumgradw", "get_ptype", "get_simtime",
           "get_dt", "get_drho", "get_dpccorr", "makedf", "writecsv", "writevtk", "collect_vtk", "VtkWriter", "create_vtp_file",
           "paraview_collection", "vtk_save", "read_vtp", "default_timelims", "build", "lib", "LIB_PATH", "GCLSPH_F64", "GCLSPH_F32",
           "GCLSPH_F32SPLIT"]
