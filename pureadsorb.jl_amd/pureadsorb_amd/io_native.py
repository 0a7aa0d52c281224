"""The input-file readers of the C ABI (pa_read_cif, pa_read_forcefield, pa_read_guest): the
same files and results as `read_cif`, `read_forcefield` and `read_guest` (src/structure.jl:37-69,
src/forcefield.jl:54-101), parsed in C++ for hosts that do not run Julia or Python's YAML.
Host only: no device is touched."""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

from . import _lib
from .forcefield import ForceField, Guest
from .structure import Framework


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def _strings(getter, handle, *extra) -> list:
    need = C.c_size_t(0)
    _lib.check(getter(handle, *extra, None, 0, C.byref(need)))
    buf = C.create_string_buffer(max(1, need.value))
    _lib.check(getter(handle, *extra, buf, len(buf), C.byref(need)))
    raw = buf.raw[:need.value]
    return [s.decode() for s in raw.split(b"\0")[:-1]] if need.value else []


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def read_cif_native(path) -> Framework:
    """Framework of a P1 CIF file through pa_read_cif."""
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.pa_read_cif(_path(path), C.byref(h)))
    try:
        n = C.c_int32()
        cell = np.zeros(9)
        _lib.check(L.pa_cif_info(h, C.byref(n), _ptr(cell)))
        frac = np.zeros((n.value, 3))
        q = np.zeros(n.value)
        _lib.check(L.pa_cif_atoms(h, _ptr(frac), _ptr(q)))
        labels = _strings(L.pa_cif_strings, h, C.c_int32(0))
        symbols = _strings(L.pa_cif_strings, h, C.c_int32(1))
    finally:
        L.pa_cif_destroy(h)
    return Framework(cell.reshape(3, 3, order="F"), frac, labels, symbols, q)


class _FFHandle:
    """Owns a pa_ff_file; kept on the ForceField so read_guest_native can look types up."""

    def __init__(self, path):
        L = _lib.lib()
        self.h = C.c_void_p()
        _lib.check(L.pa_read_forcefield(_path(path), C.byref(self.h)))
        weakref.finalize(self, L.pa_ff_destroy, self.h.value)

    def typeindex(self, name: str) -> int:
        i = C.c_int32()
        _lib.check(_lib.lib().pa_ff_typeindex(self.h, name.encode(), C.byref(i)))
        return i.value


def read_forcefield_native(path) -> ForceField:
    """ForceField of a kUPS-style YAML file through pa_read_forcefield (Lorentz-Berthelot tables
    from pa_ff_params)."""
    L = _lib.lib()
    fh = _FFHandle(path)
    n, cutoff, tail = C.c_int32(), C.c_double(), C.c_int32()
    _lib.check(L.pa_ff_info(fh.h, C.byref(n), C.byref(cutoff), C.byref(tail)))
    S = np.zeros((n.value, n.value))
    E = np.zeros((n.value, n.value))
    _lib.check(L.pa_ff_params(fh.h, None, None, _ptr(S), _ptr(E)))  # symmetric: order-free
    ff = ForceField(_strings(L.pa_ff_names, fh.h), S, E, cutoff.value, bool(tail.value))
    ff._native = fh
    return ff


def read_guest_native(path, ff: ForceField) -> Guest:
    """Guest of a kUPS-style YAML file through pa_read_guest; `ff` comes from
    read_forcefield_native (its C handle resolves the site symbols)."""
    fh = getattr(ff, "_native", None)
    if fh is None:
        raise _lib.ArgumentError("read_guest_native needs a ForceField from read_forcefield_native")
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.pa_read_guest(_path(path), fh.h, C.byref(h)))
    try:
        n = C.c_int32()
        tc, pc, om = C.c_double(), C.c_double(), C.c_double()
        _lib.check(L.pa_guest_info(h, C.byref(n), C.byref(tc), C.byref(pc), C.byref(om)))
        sites = np.zeros((n.value, 3))
        types = np.zeros(n.value, dtype=np.int32)
        q = np.zeros(n.value)
        _lib.check(L.pa_guest_sites(h, _ptr(sites), _ptr(types), _ptr(q)))
    finally:
        L.pa_guest_destroy(h)
    return Guest(sites, types, q, tc.value, pc.value, om.value)


def cif_types_native(path, ff: ForceField) -> np.ndarray:
    """pa_framework.types of a CIF file: typeindex(ff, symbol * "_") per atom (pa_cif_types)."""
    fh = getattr(ff, "_native", None)
    if fh is None:
        raise _lib.ArgumentError("cif_types_native needs a ForceField from read_forcefield_native")
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.pa_read_cif(_path(path), C.byref(h)))
    try:
        n = C.c_int32()
        _lib.check(L.pa_cif_info(h, C.byref(n), None))
        t = np.zeros(n.value, dtype=np.int32)
        _lib.check(L.pa_cif_types(h, fh.h, _ptr(t)))
    finally:
        L.pa_cif_destroy(h)
    return t
