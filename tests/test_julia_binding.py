"""The Julia binding (pureadsorb.jl_amd/julia/PureAdsorbMI355X.jl) cannot run here (no Julia), so
its C-structure mirrors and ccall targets are checked against the ones the ctypes layer uses and
the library exports: every `struct Pa*` has the ctypes structure's fields in the same order with
the same C types, and every `ccall((:pa_..., LIB), ...)` names an exported symbol."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from pureadsorb_amd import _lib

JL = os.path.join(ROOT, "pureadsorb.jl_amd", "julia", "PureAdsorbMI355X.jl")

_SCALAR = {"Float64": C.c_double, "Int32": C.c_int32, "Int64": C.c_int64, "UInt64": C.c_uint64}


def _ctype(jt: str):
    jt = jt.strip()
    if jt.startswith("Ptr{"):
        return C.c_void_p
    m = re.fullmatch(r"NTuple\{(\d+),\s*(\w+)\}", jt)
    if m:
        return _SCALAR[m.group(2)] * int(m.group(1))
    return _SCALAR[jt]


def _julia_structs():
    text = open(JL).read()
    out = {}
    for m in re.finditer(r"^struct (Pa\w+)\n(.*?)^end", text, flags=re.S | re.M):
        fields = []
        for part in re.split(r"[;\n]", m.group(2)):
            part = part.strip()
            if not part:
                continue
            name, jt = part.split("::")
            fields.append((name.strip(), _ctype(jt)))
        out[m.group(1)] = fields
    return out


@pytest.mark.parametrize("name", ["PaFramework", "PaForcefield", "PaGuest", "PaEwald", "PaWidomResult",
                                  "PaBatchTables", "PaGcmcParams", "PaGcmcResult", "PaNvtParams", "PaNvtResult"])
def test_julia_struct_matches_the_ctypes_layout(name):
    jl = _julia_structs()[name]
    py = getattr(_lib, name)._fields_
    assert [f for f, _ in jl] == [f for f, _ in py]
    for (f, a), (_, b) in zip(jl, py):
        assert C.sizeof(a) == C.sizeof(b) and C.alignment(a) == C.alignment(b), (name, f)


def test_julia_ccalls_name_exported_symbols():
    names = set(re.findall(r"ccall\(\(:(pa_\w+), LIB\)", open(JL).read()))
    assert names and names <= set(_lib.EXPORTS), names - set(_lib.EXPORTS)
