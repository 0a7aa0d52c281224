"""FrameworkBatch: the device-resident, deduplicated tables of a set of frameworks
(src/batch.jl:72-432), built by the library's `pa_batch_create`."""
from __future__ import annotations

import ctypes
import ctypes as C

import numpy as np

from . import _lib
from .forcefield import EwaldParams, ForceField, Guest, typeindex
from .structure import Framework


def _ptr(a: np.ndarray):
    return a.ctypes.data if a.size else None


class _GuestArgs:
    """Keeps the numpy buffers of a PaGuest alive."""

    def __init__(self, guest: Guest):
        self.sites = np.ascontiguousarray(guest.sites, dtype=np.float64)
        self.types = np.ascontiguousarray(guest.types, dtype=np.int32)
        self.charges = np.ascontiguousarray(guest.charges, dtype=np.float64)
        self.s = _lib.PaGuest(len(self.charges), _ptr(self.sites), _ptr(self.types), _ptr(self.charges))


class FrameworkBatch:
    """`FrameworkBatch(fws, ff, guest, ewald; fullk = false)` (src/batch.jl:237).

    Framework atoms map to LJ types by `symbol + "_"` as in the reference (src/batch.jl:309).
    With ``host_only=True`` the tables are built and validated without touching a device
    (for CPU-side checks); such a batch cannot run `widom`.
    """

    def __init__(self, fws, ff: ForceField, guest: Guest, ewald: EwaldParams, *, fullk: bool = False,
                 host_only: bool = False):
        fws = list(fws)
        if not fws:
            raise _lib.ArgumentError("FrameworkBatch needs at least one framework")
        L = _lib.lib()
        self.nsys = len(fws)
        self.fullk = bool(fullk)
        self.cutoff = float(ff.cutoff)
        self.ntypes = len(ff.names)
        self.type_names = list(ff.names)
        self.ewald_cutoff = float(ewald.cutoff)
        self.guest_sites_orig = np.array(guest.sites, dtype=np.float64)
        self.guest_types_orig = np.array(guest.types, dtype=np.int32)
        self.guest_charges_orig = np.array(guest.charges, dtype=np.float64)
        keep = []
        arr = (_lib.PaFramework * self.nsys)()
        cache: dict[int, tuple] = {}
        for s, fw in enumerate(fws):
            if not isinstance(fw, Framework):
                raise _lib.ArgumentError(f"framework {s + 1} is not a Framework")
            key = id(fw)
            if key not in cache:
                frac = np.ascontiguousarray(fw.frac, dtype=np.float64)
                types = np.ascontiguousarray([typeindex(ff, sym + "_") for sym in fw.symbols], dtype=np.int32)
                q = np.ascontiguousarray(fw.charges, dtype=np.float64)
                cache[key] = (frac, types, q)
                keep.append(cache[key])
            frac, types, q = cache[key]
            f = arr[s]
            f.cell = (C.c_double * 9)(*np.asarray(fw.cell, dtype=np.float64).T.reshape(9))
            f.natoms = len(q)
            f.frac, f.types, f.charges = _ptr(frac), _ptr(types), _ptr(q)
            f.replication = (C.c_int32 * 3)(*fw.replication)
        sig = np.ascontiguousarray(np.asarray(ff.sigma, dtype=np.float64).T.reshape(-1))
        eps = np.ascontiguousarray(np.asarray(ff.epsilon, dtype=np.float64).T.reshape(-1))
        pff = _lib.PaForcefield(ff.ntypes, _ptr(sig), _ptr(eps), float(ff.cutoff), int(bool(ff.tail)))
        g = _GuestArgs(guest)
        pew = _lib.PaEwald(float(ewald.cutoff), float(ewald.precision))
        flags = (_lib.PA_BATCH_FULLK if fullk else 0) | (_lib.PA_BATCH_HOST_ONLY if host_only else 0)
        h = C.c_void_p()
        _lib.check(L.pa_batch_create(arr, self.nsys, C.byref(pff), C.byref(g.s), C.byref(pew), flags, C.byref(h)))
        self._h = h
        self._list_capacity = 0
        self.host_only = bool(host_only)
        self._read_info()

    @classmethod
    def from_tables(cls, tables: dict, guest: Guest) -> "FrameworkBatch":
        """Takes over a FrameworkBatch the reference has built (src/batch.jl:72-104), given as a
        dict of its arrays with 0-based indices: framework_of (nsys), atom_offsets, positions
        (natoms, 3) Å, types (compact), charges, cells (nfw, 3, 3) with lattice vectors as
        columns, alphas, k_offsets, ks (nk, 3), kprefactor, Shost (complex), constant_offset,
        self_term_halfrange, sigma / epsilon (compact ntypes x ntypes), compact_to_orig, bs,
        kmin (nfw, nsites, ntypes), cutoff, ewald_cutoff, fullk. `guest` carries force-field type
        indices. The batch runs `widom`; gradients and SystemState need pa_batch_create."""
        t = tables
        nfw = len(t["cells"])
        arrs = {
            "framework_of": np.ascontiguousarray(t["framework_of"], dtype=np.int32),
            "atom_offsets": np.ascontiguousarray(t["atom_offsets"], dtype=np.int32),
            "positions": np.ascontiguousarray(np.asarray(t["positions"], dtype=np.float64).reshape(-1, 3)),
            "types": np.ascontiguousarray(t["types"], dtype=np.int32),
            "charges": np.ascontiguousarray(t["charges"], dtype=np.float64),
            # column-major 3x3 per framework: the transpose of each row-major matrix, flattened
            "cells": np.ascontiguousarray(np.asarray(t["cells"], dtype=np.float64).transpose(0, 2, 1).reshape(-1)),
            "alphas": np.ascontiguousarray(t["alphas"], dtype=np.float64),
            "k_offsets": np.ascontiguousarray(t["k_offsets"], dtype=np.int32),
            "ks": np.ascontiguousarray(np.asarray(t["ks"], dtype=np.float64).reshape(-1, 3)),
            "kprefactor": np.ascontiguousarray(t["kprefactor"], dtype=np.float64),
            "Shost": np.ascontiguousarray(np.stack([np.real(t["Shost"]), np.imag(t["Shost"])], axis=-1)
                                          .astype(np.float64).reshape(-1, 2)),
            "constant_offset": np.ascontiguousarray(t["constant_offset"], dtype=np.float64),
            "self_term_halfrange": np.ascontiguousarray(t["self_term_halfrange"], dtype=np.float64),
            "sigma": np.ascontiguousarray(np.asarray(t["sigma"], dtype=np.float64).T.reshape(-1)),
            "epsilon": np.ascontiguousarray(np.asarray(t["epsilon"], dtype=np.float64).T.reshape(-1)),
            "compact_to_orig": np.ascontiguousarray(t["compact_to_orig"], dtype=np.int32),
            "bs": np.ascontiguousarray(t["bs"], dtype=np.float64),
            "kmin": np.ascontiguousarray(np.asarray(t["kmin"], dtype=np.float64).reshape(-1)),
        }
        pt = _lib.PaBatchTables()
        pt.nframeworks = nfw
        pt.nsys = len(arrs["framework_of"])
        for k, a in arrs.items():
            setattr(pt, k, _ptr(a))
        pt.ntypes = len(arrs["compact_to_orig"])
        pt.cutoff = float(t["cutoff"])
        pt.ewald_cutoff = float(t["ewald_cutoff"])
        pt.fullk = int(bool(t.get("fullk", False)))
        g = _GuestArgs(guest)
        h = C.c_void_p()
        _lib.check(_lib.lib().pa_batch_from_tables(C.byref(pt), C.byref(g.s), C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self._list_capacity = 0
        self.host_only = False
        self.nsys = pt.nsys
        self.fullk = bool(pt.fullk)
        self.cutoff = pt.cutoff
        self.ewald_cutoff = pt.ewald_cutoff
        self.ntypes = pt.ntypes
        self.type_names = None
        self.guest_sites_orig = np.array(guest.sites, dtype=np.float64)
        self.guest_types_orig = np.array(guest.types, dtype=np.int32)
        self.guest_charges_orig = np.array(guest.charges, dtype=np.float64)
        self._read_info()
        return self

    def _read_info(self):
        L = _lib.lib()
        h = self._h
        nfw, nsys = C.c_int32(), C.c_int32()
        _lib.check(L.pa_batch_nframeworks(h, C.byref(nfw), C.byref(nsys)))
        self.nframeworks = nfw.value
        fo = np.zeros(self.nsys, dtype=np.int32)
        _lib.check(L.pa_batch_framework_of(h, fo.ctypes.data))
        self.framework_of = fo
        infos = []
        for u in range(self.nframeworks):
            info = _lib.PaFrameworkInfo()
            _lib.check(L.pa_batch_framework_info(h, u, C.byref(info)))
            infos.append(info)
        self.natoms = np.array([i.natoms for i in infos])
        self.nk = np.array([i.nk for i in infos])
        self.nk_full = np.array([i.nk_full for i in infos])
        self.volumes = np.array([i.volume for i in infos])
        self.alphas = np.array([i.alpha for i in infos])
        self.constant_offset = np.array([i.constant_offset for i in infos])
        self.self_term_halfrange = np.array([i.self_term_halfrange for i in infos])
        self.hardcore_bound = np.array([i.hardcore_bound for i in infos])

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().pa_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise _lib.ArgumentError("batch is closed")
        return self._h

    @property
    def bin_scale(self) -> float:
        """Mixed-path bin edge multiplier (0: tuned by the first large widom call)."""
        k = ctypes.c_double()
        _lib.check(_lib.lib().pa_batch_bin_scale(self.handle, ctypes.byref(k)))
        return k.value

    @bin_scale.setter
    def bin_scale(self, kappa: float):
        _lib.check(_lib.lib().pa_batch_set_bin_scale(self.handle, float(kappa)))

    @property
    def list_capacity(self) -> int:
        """Fixed mixed-path neighbour-list capacity (entries per pass); 0 sizes it per call."""
        return self._list_capacity

    @list_capacity.setter
    def list_capacity(self, capacity: int):
        _lib.check(_lib.lib().pa_batch_set_list_capacity(self.handle, int(capacity)))
        self._list_capacity = int(capacity)

    def list_stats(self) -> tuple[int, int]:
        """(capacity, largest number of list passes of any bin) of the last mixed-path call."""
        cap, passes = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.lib().pa_batch_list_stats(self.handle, ctypes.byref(cap), ctypes.byref(passes)))
        return cap.value, passes.value


def nframeworks(batch: FrameworkBatch) -> int:
    return batch.nframeworks
