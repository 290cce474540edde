"""The C-ABI library loads and exports every entry point of include/gclsph.h;
without a GPU every object constructor fails loudly (there is no CPU path)."""
import os
import re
import subprocess

import numpy as np
import pytest
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gclsph.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(g\w+|sph_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("gcl_create", "gcl_update", "gcl_partialupdate", "gcl_get_pairs", "sph_create", "sph_stepsolve",
              "sph_get_field", "gclsph_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(G):
    from gclsph import _lib

    lib = G.lib()
    names = declared_functions()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    assert set(names) <= exported


def test_library_is_built_for_gfx950(G):
    from gclsph import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def _has_gpu():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_no_cpu_fallback(G):
    with pytest.raises(RuntimeError, match="no HIP device"):
        G.GPUCellList((np.zeros(3), np.arange(3.0)), (0.5, 0.5), 0.5)


def test_argument_errors_before_device(G):
    with pytest.raises(ValueError, match="length of points not equal"):
        G.GPUCellList((np.zeros(3), np.arange(4.0)), (0.5, 0.5), 0.5)
    with pytest.raises(ValueError, match="wrong dimention"):
        G.GPUCellList((np.zeros(3),), (0.5,), 0.5)


def test_noalloc_rejects_oversized_inputs_before_touching_a_device(G):
    """gcl_neighborlist_noalloc checks its fixed maxima on the host first, so
    these errors need no GPU."""
    # stand-in device addresses: never dereferenced, the checks come first
    with pytest.raises(ValueError, match="NOALLOC_MAX_POINTS"):
        G.neighborlist_noalloc([8, 16], G.api.NOALLOC_MAX_POINTS + 1, 0.1, 0.1, (0, 0), (1, 1), 32, 8)
    with pytest.raises(ValueError, match="NOALLOC_MAX_CELLS"):
        G.neighborlist_noalloc([8, 16], 1000, 1e-4, 1e-4, (0, 0), (1, 1), 32, 8)
    with pytest.raises(ValueError, match="dimention"):
        G.neighborlist_noalloc([8], 10, 0.1, 0.1, (0,), (1,), 32, 8)
    with pytest.raises(ValueError, match="8-byte aligned"):
        G.neighborlist_noalloc([8, 16], 10, 0.1, 0.1, (0, 0), (1, 1), 4 + 8, 8)
