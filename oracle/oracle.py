"""ctypes front end of the FP64 C oracle (``oracle.c``) -- TEST INFRASTRUCTURE ONLY.

Each wrapper names the reference function it restates (file:line into PureAdsorb.jl).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

KB = 1.38064852e-23 / 1.6021766208e-19  # src/constants.jl:3
KE = 14.399645351950548  # src/constants.jl:4

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile ``liboracle.so`` with the committed Makefile (gcc, no GPU code)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.or_splitmix64.restype = C.c_uint64
    L.or_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
    L.or_mulhi64.restype = C.c_uint64
    L.or_mulhi64.argtypes = [C.c_uint64, C.c_uint64]
    L.or_unit_f64.restype = C.c_double
    L.or_unit_f64.argtypes = [C.c_uint64]
    L.or_unit_f32.restype = C.c_float
    L.or_unit_f32.argtypes = [C.c_uint64]
    L.or_chain_words.restype = None
    L.or_chain_words.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, _up]
    L.or_rand_range.restype = C.c_int64
    L.or_rand_range.argtypes = [C.c_uint64, C.c_int64]
    L.or_sincospi.restype = None
    L.or_sincospi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_shoemake.restype = None
    L.or_shoemake.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
    L.or_pose.restype = None
    L.or_pose.argtypes = [C.c_uint64, C.c_int64, C.c_int64, _dp, _dp, _dp]
    L.or_cell_matrix.restype = C.c_int
    L.or_cell_matrix.argtypes = [C.c_double] * 6 + [_dp]
    L.or_inv3.restype = None
    L.or_inv3.argtypes = [_dp, _dp]
    L.or_det3.restype = C.c_double
    L.or_det3.argtypes = [_dp]
    L.or_perpendicular_lengths.restype = None
    L.or_perpendicular_lengths.argtypes = [_dp, _dp]
    L.or_minimum_image.restype = None
    L.or_minimum_image.argtypes = [_dp, _dp, _dp, _dp]
    L.or_rotate.restype = None
    L.or_rotate.argtypes = [_dp, _dp, _dp]
    for name in ("or_normals", "or_quats"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, _dp]
    L.or_quat_pow.restype = None
    L.or_quat_pow.argtypes = [_dp, C.c_double, _dp]
    L.or_quat_mul.restype = None
    L.or_quat_mul.argtypes = [_dp, _dp, _dp]
    L.or_erfc_dev.restype = C.c_double
    L.or_erfc_dev.argtypes = [C.c_double]
    L.or_erfc_coef.restype = C.c_double
    L.or_erfc_coef.argtypes = [C.c_int, C.c_int]
    L.or_pair_erfc_f64.restype = C.c_double
    L.or_pair_erfc_f64.argtypes = [C.c_double]
    L.or_pair_erfc_f32.restype = C.c_float
    L.or_pair_erfc_f32.argtypes = [C.c_float]
    L.or_ewald_alpha.restype = C.c_double
    L.or_ewald_alpha.argtypes = [C.c_double, C.c_double]
    L.or_ewald_kmax.restype = C.c_double
    L.or_ewald_kmax.argtypes = [C.c_double, C.c_double]
    L.or_pk.restype = C.c_double
    L.or_pk.argtypes = [C.c_double, C.c_double, C.c_double]
    L.or_kvectors.restype = C.c_int
    L.or_kvectors.argtypes = [_dp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_structure_factor.restype = None
    L.or_structure_factor.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]
    L.or_ewald_energy.restype = C.c_double
    L.or_ewald_energy.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_double, C.c_double, C.c_int, _dp, _dp]
    L.or_tail_delta.restype = C.c_double
    L.or_tail_delta.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _lp, _lp, C.c_double]
    L.or_self_orientation.restype = None
    L.or_self_orientation.argtypes = [C.c_int, _dp]
    L.or_system_new.restype = C.c_void_p
    L.or_system_new.argtypes = [
        _dp, C.c_int, _dp, _ip, _dp, _ip, C.c_int, _dp, _dp, C.c_double, C.c_int,
        C.c_int, _dp, _ip, _dp, C.c_double, C.c_double,
    ]
    L.or_system_free.restype = None
    L.or_system_free.argtypes = [C.c_void_p]
    L.or_system_status.restype = C.c_int
    L.or_system_status.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    L.or_system_scalars.restype = None
    L.or_system_scalars.argtypes = [C.c_void_p, _dp]
    L.or_system_ktable.restype = None
    L.or_system_ktable.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _ip]
    L.or_insertion_energy.restype = C.c_double
    L.or_insertion_energy.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_int64)]
    L.or_insertion_energy_reference.restype = C.c_double
    L.or_insertion_energy_reference.argtypes = [C.c_void_p, _dp, _dp]
    L.or_energies.restype = None
    L.or_energies.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_void_p, C.c_void_p, C.c_int]
    L.or_default_run.restype = C.c_int64
    L.or_default_run.argtypes = [C.c_int64, C.c_int64]
    L.or_system_counts.restype = None
    L.or_system_counts.argtypes = [C.c_int64, C.c_int64, C.c_int64, _lp]
    L.or_boltzmann_weight.restype = None
    L.or_boltzmann_weight.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_reduce.restype = C.c_int
    L.or_reduce.argtypes = [C.c_int, _dp, _dp, _lp, C.c_double, C.c_double, _dp]
    L.or_widom.restype = C.c_int
    L.or_widom.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_uint64, C.c_int64, C.c_int, _dp, _dp, C.c_int]
    L.or_theta_f64.restype = C.c_double
    L.or_theta_f64.argtypes = []
    L.or_pair_energy.restype = C.c_double
    L.or_pair_energy.argtypes = [C.c_double] * 5
    L.or_find_r0.restype = C.c_double
    L.or_find_r0.argtypes = [C.c_double] * 5
    L.or_find_rho2.restype = C.c_double
    L.or_find_rho2.argtypes = [C.c_double] * 5
    L.or_hardcore_bound.restype = C.c_double
    L.or_hardcore_bound.argtypes = [C.c_void_p]
    L.or_rejection_tables.restype = C.c_double
    L.or_rejection_tables.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    L.or_hardcore_flags.restype = None
    L.or_hardcore_flags.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_void_p, C.c_int]
    L.or_energies_twophase.restype = None
    L.or_energies_twophase.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, _dp, _dp, C.c_void_p,
                                       C.c_int]
    L.or_host_guest.restype = C.c_double
    L.or_host_guest.argtypes = [C.c_void_p, _dp, _dp]
    L.or_guest_self.restype = C.c_double
    L.or_guest_self.argtypes = [C.c_void_p]
    L.or_guest_tail.restype = C.c_double
    L.or_guest_tail.argtypes = [C.c_void_p, C.c_int64]
    L.or_net.restype = C.c_double
    L.or_net.argtypes = [C.c_void_p, C.c_int64]
    L.or_exchange_constant.restype = C.c_double
    L.or_exchange_constant.argtypes = [C.c_void_p, C.c_int64]
    L.or_mc_new.restype = C.c_void_p
    L.or_mc_new.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p]
    L.or_mc_free.restype = None
    L.or_mc_free.argtypes = [C.c_void_p]
    L.or_mc_total_energy.restype = C.c_double
    L.or_mc_total_energy.argtypes = [C.c_void_p]
    L.or_mc_attempt.restype = C.c_int
    L.or_mc_attempt.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.or_mc_cycle.restype = C.c_int64
    L.or_mc_cycle.argtypes = [C.c_void_p, _dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double]
    L.or_mc_scalars.restype = None
    L.or_mc_scalars.argtypes = [C.c_void_p, _dp]
    L.or_mc_counts.restype = None
    L.or_mc_counts.argtypes = [C.c_void_p, _ip, _ip]
    L.or_mc_guests.restype = None
    L.or_mc_guests.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_mc_sk.restype = None
    L.or_mc_sk.argtypes = [C.c_void_p, _dp, _dp]
    L.or_nvt_const_term.restype = C.c_double
    L.or_nvt_const_term.argtypes = [C.c_void_p, C.c_int64]
    L.or_chain_insertion.restype = C.c_double
    L.or_chain_insertion.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    L.or_chain_energies.restype = None
    L.or_chain_energies.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_double, _dp, C.c_int]
    L.or_mc_audit.restype = C.c_double
    L.or_mc_audit.argtypes = [C.c_void_p]
    L.or_widom_grad.restype = C.c_int
    L.or_widom_grad.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_uint64, C.c_int64, C.c_int, _dp, _dp, _dp,
                                _dp, _dp, _dp, _dp]
    _lib = L
    return L


# --- scalar helpers -------------------------------------------------------------------------
def theta_f64() -> float:
    """theta_F(Float64) (src/reject.jl:11-19)."""
    return lib().or_theta_f64()


def pair_energy(r, sigma, eps, K, alpha) -> float:
    """src/reject.jl:22"""
    return lib().or_pair_energy(r, sigma, eps, K, alpha)


def find_r0(sigma, eps, K, alpha, r_lj) -> float:
    """src/reject.jl:45-69"""
    return lib().or_find_r0(sigma, eps, K, alpha, r_lj)


def find_rho2(sigma, eps, kmin, margin, r_lj) -> float:
    """src/reject.jl:191-209"""
    return lib().or_find_rho2(sigma, eps, kmin, margin, r_lj)

def splitmix64(seed: int, index: int) -> int:
    """src/state.jl:225-230"""
    return lib().or_splitmix64(seed & (2**64 - 1), index & (2**64 - 1))


def mulhi64(a: int, b: int) -> int:
    """src/rng.jl:14-24"""
    return lib().or_mulhi64(a, b)


def chain_words(chain_seed: int, cycle: int, move: int, n: int) -> np.ndarray:
    """Words 0..n-1 of ChainRNG(chain_seed, cycle, move) (src/rng.jl:43-61)."""
    out = np.zeros(n, dtype=np.uint64)
    lib().or_chain_words(chain_seed & (2**64 - 1), cycle & (2**64 - 1), move & (2**64 - 1), n, out)
    return out


def unit_f64(raw: int) -> float:
    return lib().or_unit_f64(raw)


def unit_f32(raw: int) -> float:
    return lib().or_unit_f32(raw)


def rand_range(raw: int, n: int) -> int:
    return lib().or_rand_range(raw, n)


def sincospi(x: float) -> tuple[float, float]:
    s, c = C.c_double(), C.c_double()
    lib().or_sincospi(x, C.byref(s), C.byref(c))
    return s.value, c.value


def shoemake(u1: float, u2: float, u3: float) -> np.ndarray:
    """src/rng.jl:126-129"""
    q = np.zeros(4)
    lib().or_shoemake(u1, u2, u3, q)
    return q


def normals(chain_seed: int, cycle: int, move: int, n: int) -> np.ndarray:
    """n successive rand_normal draws of one ChainRNG stream (src/rng.jl:109-141)."""
    out = np.zeros(max(int(n), 1))
    lib().or_normals(chain_seed & (2**64 - 1), cycle & (2**64 - 1), move & (2**64 - 1), int(n), out)
    return out[:int(n)]


def quaternions(chain_seed: int, cycle: int, move: int, n: int) -> np.ndarray:
    """n successive rand_quaternion draws (x, y, z, w) of one ChainRNG stream."""
    out = np.zeros((max(int(n), 1), 4))
    lib().or_quats(chain_seed & (2**64 - 1), cycle & (2**64 - 1), move & (2**64 - 1), int(n), out)
    return out[:int(n)]


def quaternion_pow(q, t: float) -> np.ndarray:
    """quaternion_pow (src/moves.jl:99-107)."""
    out = np.zeros(4)
    lib().or_quat_pow(np.ascontiguousarray(q, dtype=np.float64), float(t), out)
    return out


def qmul(q2, q1) -> np.ndarray:
    """qmul (src/moves.jl:81-88): apply q1, then q2."""
    out = np.zeros(4)
    lib().or_quat_mul(np.ascontiguousarray(q2, dtype=np.float64), np.ascontiguousarray(q1, dtype=np.float64), out)
    return out


def rotate(q, v) -> np.ndarray:
    """rotate (src/energy.jl:13-16)."""
    out = np.zeros(3)
    lib().or_rotate(np.ascontiguousarray(q, dtype=np.float64), np.ascontiguousarray(v, dtype=np.float64), out)
    return out


def pose(seed: int, sys: int, j: int, A: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """propose_reinsertion keyed ChainRNG(splitmix64(seed, sys+1), j+1, 0) (src/moves.jl:154-161)."""
    pos, q = np.zeros(3), np.zeros(4)
    lib().or_pose(seed & (2**64 - 1), sys, j, _colmajor(A), pos, q)
    return pos, q


def _colmajor(A) -> np.ndarray:
    A = np.asarray(A, dtype=np.float64)
    return np.ascontiguousarray(A.reshape(3, 3).T.reshape(9)) if A.shape == (3, 3) else np.ascontiguousarray(A)


def cell_matrix(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """src/cell.jl:2-12; returns the 3x3 matrix (lattice vectors as columns)."""
    A = np.zeros(9)
    st = lib().or_cell_matrix(a, b, c, alpha, beta, gamma, A)
    if st != 0:
        raise ValueError(f"degenerate cell (status {st})")
    return A.reshape(3, 3, order="F")


def erfc_dev(x: float) -> float:
    return lib().or_erfc_dev(x)


def erfc_coefficients(which: str) -> np.ndarray:
    """The oracle's coefficient literals: 'erfc_dev' (28), 'pair64' (17), 'pair32' (8)."""
    w, n = {"erfc_dev": (0, 28), "pair64": (1, 17), "pair32": (2, 8)}[which]
    return np.array([lib().or_erfc_coef(w, j) for j in range(n)])


def pair_erfc_f64(x: float) -> float:
    return lib().or_pair_erfc_f64(x)


def pair_erfc_f32(x: float) -> float:
    return lib().or_pair_erfc_f32(x)


def ewald_alpha(cutoff: float, precision: float) -> float:
    return lib().or_ewald_alpha(cutoff, precision)


def ewald_kmax(alpha: float, precision: float) -> float:
    return lib().or_ewald_kmax(alpha, precision)


def kvectors(A, kmax: float):
    """src/ewald.jl:142-157: (ks (n,3), weights (n,), coeffs (n,3))."""
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.reshape(9))
    n = lib().or_kvectors(Acm, kmax, 0, None, None, None)
    ks = np.zeros(3 * max(n, 1))
    w = np.zeros(max(n, 1))
    co = np.zeros(3 * max(n, 1), dtype=np.int32)
    lib().or_kvectors(Acm, kmax, n, ks.ctypes.data, w.ctypes.data, co.ctypes.data)
    return ks[: 3 * n].reshape(n, 3), w[:n], co[: 3 * n].reshape(n, 3)


def ewald_energy(A, pos, q, mol, alpha, cutoff, ks, w) -> float:
    """src/ewald.jl:177-203"""
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.reshape(9))
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(q, dtype=np.float64)
    mol = np.ascontiguousarray(mol, dtype=np.int32)
    ks = np.ascontiguousarray(ks, dtype=np.float64).reshape(-1)
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().or_ewald_energy(Acm, len(q), pos, q, mol, alpha, cutoff, len(w), ks if len(ks) else np.zeros(1), w if len(w) else np.zeros(1))


def self_orientation(o: int) -> np.ndarray:
    q = np.zeros(4)
    lib().or_self_orientation(o, q)
    return q


def default_run(ninsert: int, nsys: int) -> int:
    return lib().or_default_run(ninsert, nsys)


def system_counts(ninsert: int, nsys: int, run: int) -> np.ndarray:
    ns = np.zeros(nsys, dtype=np.int64)
    lib().or_system_counts(ninsert, nsys, run, ns)
    return ns


def boltzmann_weight(dU: float, kT: float) -> tuple[float, float]:
    """src/widom.jl:80-84: (exp(-dU/kT), dU*w) with the product forced to 0 when w == 0."""
    w, uw = C.c_double(), C.c_double()
    lib().or_boltzmann_weight(dU, kT, C.byref(w), C.byref(uw))
    return w.value, uw.value


def reduce(sW, sUW, n, kT, V):
    """src/widom.jl:451-492. Returns (status, 8 fields)."""
    out = np.zeros(8)
    sW = np.ascontiguousarray(sW, dtype=np.float64)
    sUW = np.ascontiguousarray(sUW, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.int64)
    st = lib().or_reduce(len(sW), sW, sUW, n, kT, V, out)
    return st, out


class OracleSystem:
    """One framework system with FrameworkBatch's per-framework tables (src/batch.jl:319-408).

    ``cell`` is the 3x3 cell matrix, ``frac`` (natoms, 3), ``types`` force-field type indices
    (0-based) into the ``sigma``/``epsilon`` tables, ``guest_types`` likewise.
    """

    def __init__(self, cell, frac, types, charges, replication, sigma, epsilon, cutoff, tail,
                 guest_sites, guest_types, guest_charges, ewald_cutoff, precision=1e-6):
        L = lib()
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        Acm = np.ascontiguousarray(self.cell.T.reshape(9))
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1)
        self.natoms = len(frac) // 3
        types = np.ascontiguousarray(types, dtype=np.int32)
        charges = np.ascontiguousarray(charges, dtype=np.float64)
        rep = np.ascontiguousarray(replication, dtype=np.int32)
        sigma = np.asarray(sigma, dtype=np.float64)
        epsilon = np.asarray(epsilon, dtype=np.float64)
        nt = sigma.shape[0]
        self.ntypes = nt
        self.nsites = len(np.asarray(guest_charges).reshape(-1))
        sig = np.ascontiguousarray(sigma.T.reshape(-1))
        eps = np.ascontiguousarray(epsilon.T.reshape(-1))
        gs = np.ascontiguousarray(guest_sites, dtype=np.float64).reshape(-1)
        gt = np.ascontiguousarray(guest_types, dtype=np.int32)
        gq = np.ascontiguousarray(guest_charges, dtype=np.float64)
        if self.natoms == 0:
            frac = np.zeros(3)
            types = np.zeros(1, dtype=np.int32)
            charges = np.zeros(1)
        self._h = L.or_system_new(Acm, self.natoms, frac, types, charges, rep, nt, sig, eps, cutoff, int(bool(tail)),
                                  len(gq), gs, gt, gq, ewald_cutoff, precision)
        buf = C.create_string_buffer(512)
        st = L.or_system_status(self._h, buf, 512)
        if st != 0:
            msg = buf.value.decode()
            L.or_system_free(self._h)
            self._h = None
            raise ValueError(msg)
        sc = np.zeros(12)
        L.or_system_scalars(self._h, sc)
        (self.V, self.alpha, self.kmax, self.constant_offset, self.self_mean, self.self_halfrange,
         self.tail_term, self.self_term, self.excl_term, self.net_term) = sc[:10]
        self.nk = int(sc[10])
        self.nkfull = int(sc[11])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_system_free(self._h)
            self._h = None

    def ktable(self, full=False):
        n = self.nkfull if full else self.nk
        ks = np.zeros(3 * max(n, 1)); kp = np.zeros(max(n, 1)); sr = np.zeros(max(n, 1)); si = np.zeros(max(n, 1))
        co = np.zeros(3 * max(n, 1), dtype=np.int32)
        lib().or_system_ktable(self._h, int(full), ks, kp, sr, si, co)
        return ks[: 3 * n].reshape(n, 3), kp[:n], sr[:n] + 1j * si[:n], co[: 3 * n].reshape(n, 3)

    def insertion_energy(self, pos, q):
        """src/energy.jl:42-82 (per-site imaging); returns (E, in-cutoff pair count)."""
        npair = C.c_int64()
        e = lib().or_insertion_energy(self._h, np.ascontiguousarray(pos, dtype=np.float64),
                                      np.ascontiguousarray(q, dtype=np.float64), C.byref(npair))
        return e, npair.value

    def nvt_const_term(self, Ng):
        """insertion_constant_term (src/nvt.jl:44-66) at Ng guests."""
        return lib().or_nvt_const_term(self._h, int(Ng))

    def insertion_energy_reference(self, pos, q):
        """src/reference.jl:31-67"""
        return lib().or_insertion_energy_reference(self._h, np.ascontiguousarray(pos, dtype=np.float64),
                                                   np.ascontiguousarray(q, dtype=np.float64))

    def energies(self, seed, sysidx, j0, n, want_poses=False, want_pairs=False, nthreads=0):
        """Delta U (insertion_energy + constant_offset) for insertions j0..j0+n-1."""
        dU = np.zeros(max(n, 1))
        poses = np.zeros(7 * max(n, 1)) if want_poses else None
        pairs = np.zeros(max(n, 1), dtype=np.int64) if want_pairs else None
        lib().or_energies(self._h, seed & (2**64 - 1), sysidx, j0, n, dU,
                          poses.ctypes.data if poses is not None else None,
                          pairs.ctypes.data if pairs is not None else None, nthreads)
        out = [dU[:n]]
        if want_poses:
            out.append(poses[: 7 * n].reshape(n, 7))
        if want_pairs:
            out.append(pairs[:n])
        return out[0] if len(out) == 1 else tuple(out)

    def widom(self, T, ns, seed, sysidx, nblocks, nthreads=0):
        """One system's Widom result (src/widom.jl:290-492, FP64, no rejection).

        Returns (status, fields[8], block_sums (nblocks, 3))."""
        bs = np.zeros(3 * nblocks)
        out = np.zeros(8)
        st = lib().or_widom(self._h, T, ns, seed & (2**64 - 1), sysidx, nblocks, bs, out, nthreads)
        return st, out, bs.reshape(nblocks, 3)

    def hardcore_bound(self):
        """B_s of src/reject.jl:101-131 (exact per-atom terms)."""
        return lib().or_hardcore_bound(self._h)

    def rejection_tables(self, kT, eps_f):
        """(margin, rho2[nsites, ntypes]) of build_rejection_tables (src/widom.jl:222-253)."""
        rho2 = np.zeros(self.nsites * self.ntypes)
        margin = lib().or_rejection_tables(self._h, kT, eps_f, rho2)
        return margin, rho2.reshape(self.nsites, self.ntypes)

    def hardcore_flags(self, seed, sysidx, j0, n, rho2, nthreads=0):
        """Phase-0 flags (1 = rejected) of insertions j0..j0+n-1, brute force."""
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        r = np.ascontiguousarray(rho2, dtype=np.float64).reshape(-1)
        lib().or_hardcore_flags(self._h, seed & (2**64 - 1), sysidx, j0, n, r, flags.ctypes.data, nthreads)
        return flags[:n]

    def energies_twophase(self, seed, sysidx, j0, n, rho2, nthreads=0):
        """(dU, in-cutoff pairs) of the two-phase path: rejected insertions get dU = +inf, 0 pairs."""
        dU = np.zeros(max(n, 1))
        pairs = np.zeros(max(n, 1), dtype=np.int64)
        r = np.ascontiguousarray(rho2, dtype=np.float64).reshape(-1)
        lib().or_energies_twophase(self._h, seed & (2**64 - 1), sysidx, j0, n, r, dU, pairs.ctypes.data, nthreads)
        return dU[:n], pairs[:n]

    def widom_grad(self, T, ns, seed, sysidx, nthreads=0):
        """FP64 reweighting gradients of ln K_H and q_st over insertions 0..ns-1 (or_widom_grad).

        Returns a dict: ln_K_H, q_st, dlnkh_deps, dlnkh_dsig, dqst_deps, dqst_dsig (per force-field
        type) and dlnkh_dq, dqst_dq (per framework atom, mean over replication images)."""
        nt, na = self.ntypes, self.natoms
        out = np.zeros(2)
        lj = [np.zeros(max(nt, 1)) for _ in range(4)]
        dq = [np.zeros(max(na, 1)) for _ in range(2)]
        st = lib().or_widom_grad(self._h, T, ns, seed & (2**64 - 1), sysidx, nthreads, out, *lj, *dq)
        if st != 0:
            raise RuntimeError(f"or_widom_grad: status {st}")
        return {"ln_K_H": out[0], "q_st": out[1], "dlnkh_deps": lj[0][:nt], "dlnkh_dsig": lj[1][:nt],
                "dqst_deps": lj[2][:nt], "dqst_dsig": lj[3][:nt], "dlnkh_dq": dq[0][:na], "dqst_dq": dq[1][:na]}


PASCAL = 1.0 / (1.0e30 * 1.6021766208e-19)  # Pa -> eV/Å^3, src/constants.jl:10
KB = 1.38064852e-23 / 1.6021766208e-19


class OracleMC:
    """One FP64 Monte Carlo chain over an OracleSystem built from every k-vector (the system's
    full table): or_mc_* (the reference's move path restated serially)."""

    def __init__(self, system: "OracleSystem", capacity: int, chain_seed: int, pos=None, quat=None):
        self.system = system
        pos = np.zeros((0, 3)) if pos is None else np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        quat = np.zeros((0, 4)) if quat is None else np.ascontiguousarray(quat, dtype=np.float64).reshape(-1, 4)
        self.capacity = capacity
        self._p = pos.reshape(-1).copy()
        self._q = quat.reshape(-1).copy()
        self._h = lib().or_mc_new(system._h, capacity, chain_seed & (2**64 - 1), len(pos),
                                  self._p.ctypes.data if len(pos) else None, self._q.ctypes.data if len(pos) else None)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_mc_free(self._h)
            self._h = None

    def attempt(self, movetype, step_trans=0.0, step_rot=0.0, fugacity_pa=1.0, T=298.15):
        """(status, dU, slot): status 1 accepted, 0 rejected, -1 capacity hit."""
        dU = C.c_double(0.0)
        slot = C.c_int(0)
        st = lib().or_mc_attempt(self._h, int(movetype), float(step_trans), float(step_rot), float(fugacity_pa) * PASCAL,
                                 KB * T, C.byref(dU), C.byref(slot))
        return st, dU.value, slot.value

    def cycle(self, probs, min_cycle, step_trans, step_rot, fugacity_pa, T):
        """One GCMC cycle; probs = (translation, rotation, reinsertion, exchange)."""
        pr = np.array([probs[0], probs[1], probs[2], probs[3] / 2, probs[3] / 2], dtype=np.float64)
        pc = np.cumsum(pr / pr.sum())
        pc[-1] = 2.0
        return lib().or_mc_cycle(self._h, pc, int(min_cycle), float(step_trans), float(step_rot),
                                 float(fugacity_pa) * PASCAL, KB * T)

    def scalars(self):
        out = np.zeros(6)
        lib().or_mc_scalars(self._h, out)
        return {"occupancy": int(out[0]), "energy": out[1], "energy_abs": out[2], "counter": int(out[3]),
                "const_p": out[4], "const_q": out[5]}

    def counts(self):
        acc = np.zeros(5, dtype=np.int32)
        att = np.zeros(5, dtype=np.int32)
        lib().or_mc_counts(self._h, acc, att)
        return acc, att

    def guests(self):
        n = self.scalars()["occupancy"]
        p, q, h = np.zeros((max(n, 1), 3)), np.zeros((max(n, 1), 4)), np.zeros(max(n, 1))
        lib().or_mc_guests(self._h, p.ctypes.data, q.ctypes.data, h.ctypes.data)
        return p[:n], q[:n], h[:n]

    def sk(self):
        nk = int(lib_scalars(self.system)[11])
        re, im = np.zeros(max(nk, 1)), np.zeros(max(nk, 1))
        lib().or_mc_sk(self._h, re, im)
        return re[:nk] + 1j * im[:nk]

    def total_energy(self):
        return lib().or_mc_total_energy(self._h)

    def audit(self):
        """audit_energy!'s reset (src/audit.jl:97-165); returns recomputed - running."""
        return lib().or_mc_audit(self._h)

    def const_term(self):
        """insertion_constant_term at the chain's occupancy (src/nvt.jl:44-66)."""
        return lib().or_nvt_const_term(self.system._h, self.scalars()["occupancy"])

    def chain_insertion(self, pos, q, const_term=None):
        """widom_chain_kernel! (src/nvt.jl:85-119) at one Cartesian pose."""
        ct = self.const_term() if const_term is None else const_term
        return lib().or_chain_insertion(self._h, np.ascontiguousarray(pos, dtype=np.float64),
                                        np.ascontiguousarray(q, dtype=np.float64), float(ct))

    def chain_energies(self, widom_seed, sysidx, j0, n, const_term=None, nthreads=0):
        """Test insertions j0 .. j0+n-1 of the Widom stream of system `sysidx` into the chain."""
        ct = self.const_term() if const_term is None else const_term
        out = np.zeros(max(int(n), 1))
        lib().or_chain_energies(self._h, int(widom_seed) & (2**64 - 1), int(sysidx), int(j0), int(n), float(ct), out,
                                int(nthreads))
        return out[:int(n)]


def block_mean_sem(x, nblocks):
    """block_mean_sem (src/nvt.jl:126-140): contiguous blocks, the remainder folded into the last."""
    x = np.asarray(x, dtype=np.float64)
    ns = len(x)
    if nblocks < 2 or ns < 2 * nblocks:
        raise ValueError("block_mean_sem needs nblocks >= 2 and at least 2*nblocks samples")
    bl = ns // nblocks
    means = []
    for b in range(nblocks):
        lo, hi = b * bl, (ns if b == nblocks - 1 else (b + 1) * bl)
        s = 0.0
        for v in x[lo:hi]:
            s += v
        means.append(s / (hi - lo))
    m = sum(means) / nblocks
    v = sum((y - m) ** 2 for y in means) / (nblocks - 1)
    return m, float(np.sqrt(v / nblocks))


def fluctuation_qst(energy, loading, kT, nblocks):
    """fluctuation_qst (src/gcmc.jl:71-113): q_st = kT - cov(U, N) / var(N) from the pooled
    per-cycle samples, its delete-one-block jackknife error over `nblocks` contiguous blocks (the
    remainder in the last) and the pooled correlation corr(U, N). Returns (q_st, q_err, corr)."""
    U = np.asarray(energy, dtype=np.float64)
    N = np.asarray(loading, dtype=np.float64)
    ns = len(U)
    if len(N) != ns:
        raise ValueError("energy and loading must have the same length")
    if nblocks < 2:
        raise ValueError(f"nblocks={nblocks}: at least two blocks are needed for a jackknife error")
    if ns < 2 * nblocks:
        raise ValueError(f"{ns} cycles: need at least 2*nblocks={2 * nblocks} production cycles")
    bl = ns // nblocks
    sums = np.zeros((nblocks, 5))  # U, N, U^2, N^2, U N per block, summed in sample order
    cnt = np.zeros(nblocks)
    for b in range(nblocks):
        lo, hi = b * bl, (ns if b == nblocks - 1 else (b + 1) * bl)
        for i in range(lo, hi):
            u, n = U[i], N[i]
            sums[b] += (u, n, u * u, n * n, u * n)
        cnt[b] = hi - lo
    tot = sums.sum(axis=0)

    def qst_of(sU, sN, sUN, sN2, n):
        mU, mN = sU / n, sN / n
        return kT - (sUN / n - mU * mN) / (sN2 / n - mN * mN)

    q = qst_of(tot[0], tot[1], tot[4], tot[3], ns)
    loo = np.array([qst_of(tot[0] - sums[b, 0], tot[1] - sums[b, 1], tot[4] - sums[b, 4], tot[3] - sums[b, 3],
                           ns - cnt[b]) for b in range(nblocks)])
    qerr = math.sqrt((nblocks - 1) / nblocks * float(np.sum((loo - loo.mean()) ** 2)))
    mU, mN = tot[0] / ns, tot[1] / ns
    corr = (tot[4] / ns - mU * mN) / math.sqrt((tot[2] / ns - mU * mU) * (tot[3] / ns - mN * mN))
    return float(q), qerr, float(corr)


def nvt_run(chains, *, T, n_warmup, n_production, n_widom_per_cycle, n_audit, step_trans, step_rot,
            min_cycle_length=1, widom_seed=0, nblocks=10, nslots=None):
    """run_nvt! (src/nvt.jl:192-358) restated over OracleMC chains (one per system, system
    index = list position): nsteps = max(max occupancy, min_cycle_length) attempts per cycle,
    each chain drawing its move type uniformly among translation, rotation and reinsertion from
    its own stream (or_mc_cycle with exchange off); the audit reset every n_audit cycles; after
    every production cycle n_widom_per_cycle insertions per system from the Widom stream. Moves
    are skipped when the batch has no guest slot (nslots = total capacity, default from the
    chains). Returns per system a dict: energies, sW, sUW per cycle, dU (cycles x insertions),
    acceptance and the NVTResult fields."""
    ns = len(chains)
    kT = KB * T
    occ = [c.scalars()["occupancy"] for c in chains]
    nsteps = max(max(occ), int(min_cycle_length))
    if nslots is None:
        nslots = sum(c.capacity for c in chains)
    probs = (1.0, 1.0, 1.0, 0.0)
    ct = [c.const_term() for c in chains]
    st = np.broadcast_to(np.asarray(step_trans, dtype=np.float64), (ns,))
    sr = np.broadcast_to(np.asarray(step_rot, dtype=np.float64), (ns,))
    since = 0
    base = [None] * ns
    recs = [{"energies": [], "sW": [], "sUW": [], "dU": []} for _ in range(ns)]

    def cycle():
        nonlocal since
        if nslots > 0:
            for i, c in enumerate(chains):
                c.cycle(probs, nsteps, st[i], sr[i], 1.0, T)
        since += 1
        if since >= n_audit:
            for c in chains:
                c.audit()
            since = 0

    for _ in range(n_warmup):
        cycle()
    base = [c.counts() for c in chains]
    for cyc in range(n_production):
        cycle()
        for i, c in enumerate(chains):
            recs[i]["energies"].append(c.scalars()["energy"])
            dU = c.chain_energies(widom_seed, i, cyc * n_widom_per_cycle, n_widom_per_cycle, ct[i])
            sw = suw = 0.0
            for d in dU:
                w, uw = boltzmann_weight(d, kT)
                sw += w
                suw += uw
            recs[i]["sW"].append(sw)
            recs[i]["sUW"].append(suw)
            recs[i]["dU"].append(dU)
    out = []
    for i, c in enumerate(chains):
        r = recs[i]
        acc, att = c.counts()
        acc = acc - base[i][0]
        att = att - base[i][1]
        V = lib_scalars(c.system)[0]
        status, res = reduce(np.array(r["sW"]), np.array(r["sUW"]),
                             np.full(n_production, n_widom_per_cycle, dtype=np.int64), kT, V)
        if status != 0:
            raise ValueError(f"system {i + 1}: _reduce status {status}")
        e, ee = block_mean_sem(r["energies"], nblocks)
        out.append({"energies": np.array(r["energies"]), "sW": np.array(r["sW"]), "sUW": np.array(r["sUW"]),
                    "dU": np.array(r["dU"]), "energy": e, "energy_err": ee, "mu_ex": res[0], "mu_ex_err": res[1],
                    "K_H": res[2], "K_H_err": res[3], "q_st": res[4], "q_st_err": res[5],
                    "acceptance": tuple(float(acc[t]) / att[t] if att[t] else 0.0 for t in range(3)),
                    "attempted": att, "accepted": acc})
    return out


def lib_scalars(system):
    out = np.zeros(12)
    lib().or_system_scalars(system._h, out)
    return out
