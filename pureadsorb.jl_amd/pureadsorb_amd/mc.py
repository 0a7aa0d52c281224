"""The move-energy path on the MI355X backend: `SystemState` (src/state.jl:89-446), the NVT
and μVT moves (`mc_step!`, `mc_insert!`, `mc_delete!`, `mc_exchange!`, src/moves.jl), the
energy audit (`audit_energy!`, src/audit.jl), the GCMC driver (`run_gcmc!`, src/gcmc.jl) and the
NVT driver with Widom insertion along the chain (`run_nvt!`, `widom_chain_kernel!`, src/nvt.jl).

Every chain lives on the device; the moves run in FP64 in one launch per call."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .batch import FrameworkBatch, _GuestArgs
from .forcefield import Guest

MOVE_TRANSLATION = 1
MOVE_ROTATION = 2
MOVE_REINSERTION = 3
MOVE_INSERTION = 4
MOVE_DELETION = 5
MOVE_NAMES = ("translation", "rotation", "reinsertion", "insertion", "deletion")


def _T(T):
    return float(getattr(T, "magnitude", T))


def _vec(x, n, name):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
    if a.shape != (n,):
        raise _lib.DimensionMismatch(f"{name} must have one entry per system (nsys={n})")
    return a


class SystemState:
    """`SystemState(batch, guest, ncounts; T, seed, capacities)` (src/state.jl:328-446): one chain
    per system of `batch` (which must be built with `fullk=True`), `ncounts[n]` initial guests in
    `capacities[n]` slots."""

    def __init__(self, batch: FrameworkBatch, guest: Guest, ncounts, *, T, seed: int = 0, capacities=None):
        self.batch = batch
        self.nsys = batch.nsys
        nc = np.ascontiguousarray(np.asarray(ncounts, dtype=np.int32).reshape(-1))
        if len(nc) != self.nsys:
            raise _lib.DimensionMismatch(f"ncounts must have one entry per system (nsys={self.nsys}), got {len(nc)}")
        cap = None
        if capacities is not None:
            cap = np.ascontiguousarray(np.asarray(capacities, dtype=np.int32).reshape(-1))
            if len(cap) != self.nsys:
                raise _lib.DimensionMismatch(f"capacities must have one entry per system (nsys={self.nsys}), got {len(cap)}")
        self._g = _GuestArgs(guest)
        h = C.c_void_p()
        _lib.check(_lib.lib().pa_state_create(batch.handle, C.byref(self._g.s), nc.ctypes.data,
                                              cap.ctypes.data if cap is not None else None, _T(T),
                                              int(seed) & (2**64 - 1), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise _lib.ArgumentError("SystemState has been released")
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().pa_state_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self, n: int) -> dict:
        x = _lib.PaChainInfo()
        _lib.check(_lib.lib().pa_state_info(self.handle, int(n), C.byref(x)))
        return {"occupancy": x.occupancy, "capacity": x.capacity, "energy": x.energy,
                "energy_abs_accum": x.energy_abs_accum, "rng_counter": int(x.rng_counter),
                "accepted": list(x.accepted), "attempted": list(x.attempted), "const_p": x.const_p,
                "const_q": x.const_q}

    def nguests(self, n: int) -> int:
        return self.info(n)["occupancy"]

    def energy(self) -> np.ndarray:
        return np.array([self.info(n)["energy"] for n in range(self.nsys)])

    def guests(self, n: int):
        """(refpoints (N, 3) Å, orientations (N, 4) x y z w, cached host-guest energies (N,))."""
        occ = self.nguests(n)
        p, q, e = np.zeros((max(occ, 1), 3)), np.zeros((max(occ, 1), 4)), np.zeros(max(occ, 1))
        _lib.check(_lib.lib().pa_state_guests(self.handle, int(n), p.ctypes.data, q.ctypes.data, e.ctypes.data))
        return p[:occ], q[:occ], e[:occ]

    def sk(self, n: int) -> np.ndarray:
        nk = int(self.batch.nk[self.batch.framework_of[n]])
        s = np.zeros((max(nk, 1), 2))
        _lib.check(_lib.lib().pa_state_sk(self.handle, int(n), s.ctypes.data))
        return s[:nk, 0] + 1j * s[:nk, 1]


def mc_step(state: SystemState, movetype: int, step_trans, step_rot, *, T):
    """`mc_step!` (src/moves.jl:577-645): one attempt of `movetype` on every chain."""
    st = _vec(step_trans, state.nsys, "step_trans")
    sr = _vec(step_rot, state.nsys, "step_rot")
    _lib.check(_lib.lib().pa_mc_step(state.handle, int(movetype), st.ctypes.data, sr.ctypes.data, _T(T)))


def mc_insert(state: SystemState, fugacity, *, T):
    """`mc_insert!` (src/moves.jl:1195-1238), fugacity in Pa per system."""
    f = _vec(fugacity, state.nsys, "fugacity")
    _lib.check(_lib.lib().pa_mc_insert(state.handle, f.ctypes.data, _T(T)))


def mc_delete(state: SystemState, fugacity, *, T):
    """`mc_delete!` (src/moves.jl:1240-1290), fugacity in Pa per system."""
    f = _vec(fugacity, state.nsys, "fugacity")
    _lib.check(_lib.lib().pa_mc_delete(state.handle, f.ctypes.data, _T(T)))


def mc_exchange(rng: np.random.Generator, state: SystemState, fugacity, *, T) -> bool:
    """`mc_exchange!` (src/moves.jl:1295-1307): a fair insert/delete coin from `rng`, then that move."""
    do_insert = bool(rng.integers(2))
    (mc_insert if do_insert else mc_delete)(state, fugacity, T=T)
    return do_insert


def mc_moves(state: SystemState, movetype: int, nmoves: int, *, T, step_trans=0.0, step_rot=0.0, fugacity=None,
             trace: bool = False):
    """`nmoves` attempts of one move type on every chain in one launch. With trace=True returns
    an array (nsys, nmoves, 4): move type, slot, dU (eV), accepted."""
    st = _vec(step_trans, state.nsys, "step_trans")
    sr = _vec(step_rot, state.nsys, "step_rot")
    f = _vec(1.0 if fugacity is None else fugacity, state.nsys, "fugacity")
    tr = np.zeros((state.nsys, max(nmoves, 1), 4)) if trace else None
    _lib.check(_lib.lib().pa_mc_moves(state.handle, int(movetype), int(nmoves), st.ctypes.data, sr.ctypes.data,
                                      f.ctypes.data, _T(T), tr.ctypes.data if trace else None))
    return tr[:, :nmoves] if trace else None


def total_energy(state: SystemState) -> np.ndarray:
    """`total_energy` (src/guest.jl:293-343) of every system from its poses (FP64, host)."""
    e = np.zeros(state.nsys)
    _lib.check(_lib.lib().pa_total_energy(state.handle, e.ctypes.data))
    return e


def audit_energy(state: SystemState) -> np.ndarray:
    """`audit_energy!` (src/audit.jl:97-165) on every system, rebuilt on the device; returns
    recomputed - running (eV). A failed stage raises ArgumentError with the reference's message."""
    d = np.zeros(state.nsys)
    _lib.check(_lib.lib().pa_audit_energy(state.handle, d.ctypes.data))
    return d


def set_state_value(state: SystemState, field: str, n: int, value, index: int = 0):
    """Test-only: overwrite one element of system n's running state, for fault injection.
    field: "energy" (eV), "sk" (complex, k-vector `index`), "host_energy" (eV, live guest
    `index`) or "occupancy" (an integer; above the capacity, the state is refused by every call
    that reads its guest slots until it is set back)."""
    f = {"energy": _lib.PA_FIELD_ENERGY, "sk": _lib.PA_FIELD_SK, "host_energy": _lib.PA_FIELD_HOST_ENERGY,
         "occupancy": _lib.PA_FIELD_OCCUPANCY}.get(field)
    if f is None:
        raise _lib.ArgumentError(f"field={field!r}: use 'energy', 'sk', 'host_energy' or 'occupancy'")
    v = complex(value)
    _lib.check(_lib.lib().pa_state_set_value(state.handle, f, int(n), int(index), v.real, v.imag))


@dataclass(frozen=True)
class GCMCResult:
    """src/gcmc.jl:32-43, plus the production move counters (translation, rotation,
    reinsertion, insertion, deletion)."""
    loading: float
    loading_err: float
    energy: float
    energy_err: float
    q_st: float
    q_st_err: float
    corr_UN: float
    max_occupancy: int
    capacity: int
    ncycles: int
    accepted: tuple
    attempted: tuple


def run_gcmc(state: SystemState, *, T, n_warmup: int, n_production: int, n_audit: int, step_trans, step_rot,
             fugacity, p_translation: float = 0.25, p_rotation: float = 0.25, p_reinsertion: float = 0.0,
             p_exchange: float = 0.5, min_cycle_length: int = 1, nblocks: int = 10) -> list[GCMCResult]:
    """GCMC cycles on every chain (src/gcmc.jl:161-337): a cycle of chain n is
    max(occupancy, min_cycle_length) attempts, each a translation, rotation, reinsertion or
    exchange (a fair insertion/deletion coin) with the given probabilities, drawn per attempt from
    the chain's own counter-based stream. Returns one GCMCResult per system."""
    n = state.nsys
    f = _vec(fugacity, n, "fugacity")
    st = _vec(step_trans, n, "step_trans")
    sr = _vec(step_rot, n, "step_rot")
    p = _lib.PaGcmcParams(_T(T), int(n_warmup), int(n_production), int(n_audit), int(min_cycle_length),
                          f.ctypes.data, st.ctypes.data, sr.ctypes.data, float(p_translation), float(p_rotation),
                          float(p_reinsertion), float(p_exchange), int(nblocks))
    out = (_lib.PaGcmcResult * n)()
    _lib.check(_lib.lib().pa_gcmc_run(state.handle, C.byref(p), out))
    return [GCMCResult(r.loading, r.loading_err, r.energy, r.energy_err, r.q_st, r.q_st_err, r.corr_UN,
                       int(r.max_occupancy), int(r.capacity), int(r.ncycles), tuple(int(x) for x in r.accepted),
                       tuple(int(x) for x in r.attempted)) for r in out]


def gcmc_samples(state: SystemState, n: int, ncycles: int):
    """Per-cycle (energy, loading) production samples of system n from the last run_gcmc."""
    e, l = np.zeros(max(ncycles, 1)), np.zeros(max(ncycles, 1))
    _lib.check(_lib.lib().pa_gcmc_samples(state.handle, int(n), int(ncycles), e.ctypes.data, l.ctypes.data))
    return e[:ncycles], l[:ncycles]


def widom_chain_energies(state: SystemState, n: int, pos, quat) -> np.ndarray:
    """`widom_chain_kernel!` (src/nvt.jl:85-119) for system n (0-based): Delta U (eV) of test
    insertions at the given poses (Cartesian positions in Å, quaternions x y z w) into the
    system as it currently stands, including `insertion_constant_term` at its occupancy. The
    state is not changed."""
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    q = np.ascontiguousarray(np.asarray(quat, dtype=np.float64).reshape(-1, 4))
    if len(p) != len(q):
        raise _lib.DimensionMismatch(f"{len(p)} positions but {len(q)} orientations")
    out = np.zeros(max(len(p), 1))
    _lib.check(_lib.lib().pa_widom_chain_energies(state.handle, int(n), len(p), p.ctypes.data, q.ctypes.data,
                                                  out.ctypes.data))
    return out[:len(p)]


@dataclass(frozen=True)
class NVTResult:
    """`NVTResult` (src/nvt.jl:22-33): production energy and its cycle-block error, the Widom
    estimators from insertions into the chain's configuration (each production cycle one block),
    the production acceptance rate of translation, rotation and reinsertion."""
    energy: float
    energy_err: float
    mu_ex: float
    mu_ex_err: float
    K_H: float
    K_H_err: float
    q_st: float
    q_st_err: float
    acceptance: tuple
    ncycles: int


def run_nvt(batch: FrameworkBatch, state: SystemState, guest: Guest, ff=None, *, T, n_warmup: int,
            n_production: int, n_widom_per_cycle: int, n_audit: int, step_trans, step_rot,
            min_cycle_length: int = 1, widom_seed: int = 0, nblocks: int = 10) -> list[NVTResult]:
    """`run_nvt!(batch, state, guest, ff; T, n_warmup, n_production, n_widom_per_cycle, n_audit,
    step_trans, step_rot, min_cycle_length, widom_seed, nblocks)` (src/nvt.jl:192-358).

    A cycle is max(max occupancy, min_cycle_length) attempts on every chain, each a translation,
    rotation or reinsertion drawn uniformly from the chain's own counter-based stream (the
    reference's host `Xoshiro(seed)` move-type draw has no counterpart, so there is no `seed`
    argument); after every production cycle, n_widom_per_cycle test insertions per system whose
    poses are pa_widom's poses under `widom_seed`. `ff` is accepted for signature parity: the
    batch already holds the force field."""
    if state.batch is not batch:
        raise _lib.ArgumentError("state was not built on this batch")
    n = state.nsys
    g = _GuestArgs(guest)
    st = _vec(step_trans, n, "step_trans")
    sr = _vec(step_rot, n, "step_rot")
    p = _lib.PaNvtParams(_T(T), int(n_warmup), int(n_production), int(n_widom_per_cycle), int(n_audit),
                         int(min_cycle_length), st.ctypes.data, sr.ctypes.data, int(widom_seed) & (2**64 - 1),
                         int(nblocks))
    out = (_lib.PaNvtResult * n)()
    _lib.check(_lib.lib().pa_nvt_run(state.handle, C.byref(g.s), C.byref(p), out))
    return [NVTResult(r.energy, r.energy_err, r.mu_ex, r.mu_ex_err, r.K_H, r.K_H_err, r.q_st, r.q_st_err,
                      tuple(float(x) for x in r.acceptance), int(r.ncycles)) for r in out]


def nvt_samples(state: SystemState, n: int, ncycles: int):
    """Per-production-cycle (energy, sum of exp(-dU/kT), sum of dU exp(-dU/kT)) of system n from
    the last run_nvt."""
    e, w, uw = np.zeros(max(ncycles, 1)), np.zeros(max(ncycles, 1)), np.zeros(max(ncycles, 1))
    _lib.check(_lib.lib().pa_nvt_samples(state.handle, int(n), int(ncycles), e.ctypes.data, w.ctypes.data,
                                         uw.ctypes.data))
    return e[:ncycles], w[:ncycles], uw[:ncycles]


def nvt_last_timing(state: SystemState) -> dict:
    """Device time of the move and test-insertion launches of the last run_nvt or
    widom_chain_energies on `state`, and the host wall time of the last run_nvt."""
    t = _lib.PaNvtTiming()
    _lib.check(_lib.lib().pa_nvt_last_timing(state.handle, C.byref(t)))
    return {k: getattr(t, k) for k, _ in t._fields_}
