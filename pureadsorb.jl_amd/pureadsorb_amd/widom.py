"""`widom` on the MI355X backend (src/widom.jl:255-442) and its bookkeeping helpers."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .batch import FrameworkBatch, _GuestArgs
from .forcefield import Guest

KB = 1.38064852e-23 / 1.6021766208e-19  # eV/K, src/constants.jl:3
KE = 14.399645351950548  # eV Å / e², src/constants.jl:4
PASCAL = 1 / (1.0e30 * 1.6021766208e-19)  # Pa -> eV/Å³, src/constants.jl:10

_PRECISIONS = {"mixed": _lib.PA_PREC_MIXED, "f64": _lib.PA_PREC_F64}


@dataclass(frozen=True)
class WidomResult:
    """src/widom.jl:19-28; every field Float64. ln_K_H (ln of K_H in Å³/eV) is an addition."""
    mu_ex: float
    mu_ex_err: float
    K_H: float
    K_H_err: float
    q_st: float
    q_st_err: float
    nsamples: int
    nblocks: int
    ln_K_H: float


def sys_of_index(g: int, run: int, nsys: int) -> int:
    """1-based system of 1-based global insertion g under run-length assignment (src/widom.jl:34)."""
    return ((g - 1) // run) % nsys + 1


def default_run(ninsert: int, nsys: int) -> int:
    """src/widom.jl:39"""
    return min(max((ninsert // nsys) // 4, 1), 256)


def system_counts(ninsert: int, nsys: int, run: int) -> np.ndarray:
    """Exact per-system insertion counts (src/widom.jl:44-53)."""
    q, rem = divmod(ninsert, run)
    full, hi = divmod(q, nsys)
    ns = np.full(nsys, full * run, dtype=np.int64)
    ns[:hi] += run
    if rem > 0:
        ns[q % nsys] += rem
    return ns


def _precision(p) -> int:
    try:
        return _PRECISIONS[p]
    except KeyError:
        raise _lib.ArgumentError(f"precision={p!r}: use 'mixed' or 'f64'") from None


def widom(batch: FrameworkBatch, guest: Guest, *, T, ninsert: int, backend: str = "mi355x", seed: int = 0,
          chunk: int | None = None, nblocks: int = 10, run: int | None = None,
          precision: str = "mixed", single_phase: bool = False, no_alloc: bool = False) -> list[WidomResult]:
    """Widom test-particle insertion on the device; one WidomResult per system.

    Insertion j (0-based) of system s takes its pose from ChainRNG(splitmix64(seed, s+1), j+1, 0),
    so results are independent of `chunk` (insertions per energy launch; the library picks
    one when None) and, apart from the per-system counts, of `run`.

    The mixed path first rejects insertions whose Boltzmann weight a rigorous bound proves to be
    exactly 0.0 in Float64 (src/widom.jl:99-151, src/reject.jl); `single_phase=True` evaluates
    every insertion instead (the reference's `widom_singlephase`). Results are the same.

    `no_alloc=True` calls pa_widom_noalloc: no device allocation, every per-call buffer from one
    fixed 64 MiB device array shared by all such calls on the device (overlapping calls race),
    and no bin-scale tuning (the batch's scale as it stands).
    """
    if backend not in ("mi355x", "hip"):
        raise _lib.ArgumentError(f"backend={backend!r}: this backend runs on the MI355X ('mi355x')")
    if nblocks < 2:
        raise _lib.ArgumentError(f"nblocks={nblocks}: at least two blocks are needed for a standard error")
    if chunk is not None and chunk < 1:
        raise _lib.ArgumentError(f"chunk={chunk} must be ≥ 1")
    if run is not None and run < 1:
        # the C ABI reads run = 0 as "default"; an explicit run must be >= 1 (src/widom.jl:320)
        raise _lib.ArgumentError(f"run={run} must be ≥ 1")
    T = float(getattr(T, "magnitude", T))
    g = _GuestArgs(guest)
    out = (_lib.PaWidomResult * batch.nsys)()
    fn = _lib.lib().pa_widom_noalloc if no_alloc else _lib.lib().pa_widom
    _lib.check(fn(batch.handle, C.byref(g.s), T, int(ninsert), int(seed) & (2**64 - 1), int(nblocks),
                  0 if run is None else int(run), 0 if chunk is None else int(chunk),
                  _precision(precision) | (_lib.PA_SINGLE_PHASE if single_phase else 0), out))
    return [WidomResult(r.mu_ex, r.mu_ex_err, r.K_H, r.K_H_err, r.q_st, r.q_st_err, int(r.nsamples), int(r.nblocks),
                        r.ln_K_H) for r in out]


@dataclass(frozen=True)
class WidomGrad:
    """Gradients of ln K_H (per eV or per Å or per e) and q_st (eV per eV, Å or e) of one system.

    d*_deps / d*_dsig are indexed by force-field type (the per-type epsilon_t and sigma_t that
    Lorentz-Berthelot mixing combines); d*_dq by framework atom, each entry the derivative
    with respect to that atom's charge with its replication images moving with it, divided
    by the number of images (so a chain rule through `replicate` sums them back)."""
    result: WidomResult
    dlnkh_deps: np.ndarray
    dlnkh_dsig: np.ndarray
    dqst_deps: np.ndarray
    dqst_dsig: np.ndarray
    dlnkh_dq: np.ndarray
    dqst_dq: np.ndarray


def widom_grad(batch: FrameworkBatch, guest: Guest, *, T, ninsert: int, seed: int = 0, nblocks: int = 10,
               run: int | None = None, precision: str = "mixed") -> list[WidomGrad]:
    """`widom` plus the reweighting gradients of ln K_H and q_st over the force field's LJ
    parameters and the framework charges, from the same insertions:

        d ln K_H / dθ = -β <w ∂U/∂θ> / <w>
        d q_st / dθ   = -(<w ∂U/∂θ> - β <w U ∂U/∂θ>) / <w> - β (<w U>/<w>) <w ∂U/∂θ> / <w>

    `precision` selects the energy path whose weights are used; the gradient kernel evaluates
    its pair and k-space terms in FP32 and sums them in FP64."""
    if nblocks < 2:
        raise _lib.ArgumentError(f"nblocks={nblocks}: at least two blocks are needed for a standard error")
    if run is not None and run < 1:
        raise _lib.ArgumentError(f"run={run} must be ≥ 1")
    T = float(getattr(T, "magnitude", T))
    g = _GuestArgs(guest)
    ns, nt = batch.nsys, batch.ntypes
    nat = np.array([batch.natoms[f] for f in batch.framework_of], dtype=np.int64)
    qoff = np.concatenate([[0], np.cumsum(nat)])
    out = (_lib.PaWidomResult * ns)()
    lj = [np.zeros(max(ns * nt, 1)) for _ in range(4)]
    dq = [np.zeros(max(int(qoff[-1]), 1)) for _ in range(2)]
    _lib.check(_lib.lib().pa_widom_grad(batch.handle, C.byref(g.s), T, int(ninsert), int(seed) & (2**64 - 1),
                                        int(nblocks), 0 if run is None else int(run), _precision(precision), out,
                                        *[a.ctypes.data for a in lj], *[a.ctypes.data for a in dq]))
    res = []
    for s, r in enumerate(out):
        wr = WidomResult(r.mu_ex, r.mu_ex_err, r.K_H, r.K_H_err, r.q_st, r.q_st_err, int(r.nsamples), int(r.nblocks),
                         r.ln_K_H)
        sl = slice(s * nt, (s + 1) * nt)
        ql = slice(int(qoff[s]), int(qoff[s + 1]))
        res.append(WidomGrad(wr, lj[0][sl].copy(), lj[1][sl].copy(), lj[2][sl].copy(), lj[3][sl].copy(),
                             dq[0][ql].copy(), dq[1][ql].copy()))
    return res


def widom_energies(batch: FrameworkBatch, sys: int, seed: int, j0: int, n: int, precision: str = "mixed") -> np.ndarray:
    """Device Delta U (eV) of insertions j0..j0+n-1 of system `sys` (0-based)."""
    dU = np.zeros(max(n, 1))
    _lib.check(_lib.lib().pa_widom_energies(batch.handle, int(sys), int(seed) & (2**64 - 1), int(j0), int(n),
                                            _precision(precision), dU.ctypes.data))
    return dU[:n]


def hardcore_flags(batch: FrameworkBatch, sys: int, seed: int, T, j0: int, n: int) -> np.ndarray:
    """Phase-0 flags (1 = rejected without an energy) of insertions j0..j0+n-1 of system `sys`
    at temperature T, from the mixed path's hard-core check (src/widom.jl:99-151)."""
    T = float(getattr(T, "magnitude", T))
    flags = np.zeros(max(n, 1), dtype=np.uint8)
    _lib.check(_lib.lib().pa_hardcore_flags(batch.handle, int(sys), int(seed) & (2**64 - 1), T, int(j0), int(n),
                                            flags.ctypes.data))
    return flags[:n]


def widom_poses(batch: FrameworkBatch, sys: int, seed: int, j0: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Device-drawn poses: positions (n, 3) Å and quaternions (n, 4) in (x, y, z, w) order."""
    pos = np.zeros((max(n, 1), 3))
    quat = np.zeros((max(n, 1), 4))
    _lib.check(_lib.lib().pa_widom_poses(batch.handle, int(sys), int(seed) & (2**64 - 1), int(j0), int(n),
                                         pos.ctypes.data, quat.ctypes.data))
    return pos[:n], quat[:n]


def last_timing(batch: FrameworkBatch) -> dict:
    """Device timing of the last `widom` call on `batch` (HIP events on the library's stream)."""
    t = _lib.PaTiming()
    _lib.check(_lib.lib().pa_last_timing(batch.handle, C.byref(t)))
    return {"total_ms": t.total_ms, "energy_ms": t.energy_ms, "energy_launches": int(t.energy_launches),
            "reduce_ms": t.reduce_ms, "insertions": int(t.insertions), "binning_ms": t.binning_ms,
            "grad_ms": t.grad_ms, "grad_insertions": int(t.grad_insertions)}
