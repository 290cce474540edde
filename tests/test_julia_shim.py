"""Static checks of the Julia `ccall` binding shown in INTEGRATION.md (Julia is
not installed here, so the shim cannot run; these checks catch the mistakes a
reader of the text would otherwise have to find by hand).

* SPHProblem serves its parameters through getproperty/setproperty!
  (`prob.h` is the smoothing length, as in src/sphproblem.jl:62-112), so no
  struct field of the shim may share a name with a parameter;
* every library handle is read with `handle(x)` = getfield(x, :ptr), never as
  a property, and every struct that holds one names it `ptr`;
* every C function the shim calls is declared in include/gclsph.h, and the
  PARAMS order is the header's SPH_* parameter enum order.
"""
import os
import re

from conftest import ROOT

DOC = os.path.join(ROOT, "INTEGRATION.md")
HEADER = os.path.join(ROOT, "include", "gclsph.h")


def julia_block():
    s = open(DOC).read()
    blocks = re.findall(r"```julia\n(.*?)```", s, flags=re.S)
    assert blocks, "INTEGRATION.md holds no Julia block"
    return "\n".join(blocks)


def params(src):
    m = re.search(r"const PARAMS = \((.*?)\)\n", src, flags=re.S)
    assert m, "PARAMS tuple not found"
    return [t.strip()[1:] for t in m.group(1).replace("\n", " ").split(",") if t.strip()]


def struct_fields(src):
    out = {}
    for m in re.finditer(r"mutable struct (\w+)(?:\{[^}]*\})?\n(.*?)\n\s*(?:function|end)\b", src, flags=re.S):
        fields = re.findall(r"^\s*(\w+)::", m.group(2), flags=re.M)
        out[m.group(1)] = fields
    return out


def test_no_struct_field_is_a_parameter_name():
    src = julia_block()
    pars = set(params(src))
    assert "h" in pars and "Δt" in pars
    structs = struct_fields(src)
    assert {"GPUCellList", "SPHProblem", "SPHBatch"} <= set(structs)
    for name, fields in structs.items():
        clash = set(fields) & pars
        assert not clash, f"{name} field(s) {sorted(clash)} shadow SPHProblem parameters"


def test_handles_are_read_with_getfield():
    src = julia_block()
    assert re.search(r"^handle\(x\) = getfield\(x, :ptr\)$", src, flags=re.M)
    for name, fields in struct_fields(src).items():
        assert fields and fields[0] == "ptr", f"{name} keeps its library handle in field `ptr`"
    # no handle is read as a property (prob.ptr would go through getproperty)
    assert not re.search(r"\b\w+\.ptr\b", src)
    # the handle argument of every ccall on an object is handle(...)
    for m in re.finditer(r"ccall\(\(:(\w+), LIB\), Cint,\s*\((.*?)\),\s*(.*?)\)\)?", src, flags=re.S):
        fn, types, args = m.group(1), m.group(2), m.group(3)
        if fn in ("gclsph_last_error", "gclsph_version", "gcl_create", "sph_batch_create"):
            continue
        if types.startswith("Ptr{Cvoid}"):
            first = args.split(",")[0].strip()
            assert first.startswith("handle("), f"{fn}: handle argument is `{first}`"


def test_called_functions_are_declared_and_params_follow_the_enum():
    src = julia_block()
    hdr = open(HEADER).read()
    for fn in set(re.findall(r"ccall\(\(:(\w+), LIB\)", src)):
        assert re.search(r"\b%s\s*\(" % fn, hdr), f"{fn} is not declared in include/gclsph.h"
    enum = re.search(r"enum \{\s*(SPH_DX = 0.*?SPH_NPARAMS)", hdr, flags=re.S).group(1)
    names = [re.sub(r"/\*.*?\*/", "", t).split("=")[0].strip() for t in enum.split(",")]
    names = [n for n in names if n and n != "SPH_NPARAMS"]
    assert len(names) == len(params(src)) == 24
