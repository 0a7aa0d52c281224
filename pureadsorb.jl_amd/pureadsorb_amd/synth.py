"""Deterministic synthetic frameworks and guests for the benchmark workloads.

Every random draw comes from the same counter-based generator the device uses
(ChainRNG = SplitMix64 of a key, src/rng.jl), so a framework is a pure function of its
parameters and seed:

* atom i (0-based) of a unit cell draws its position on attempt k (1, 2, ...) from
  ChainRNG(seed, i+1, k) words 0..2 (F64 uniforms), accepted when its minimum-image distance
  to every atom already placed is at least `min_sep` (random sequential addition);
* its type from ChainRNG(seed, i+1, 0) word 0 by Lemire's map onto `types`;
* its charge from ChainRNG(seed ^ CHARGE_KEY, i+1, 0) words 0, 1 by Box-Muller (cosine branch),
  scaled by `charge_sd`; the charges are then shifted so they sum to zero.

A target count above `max_density` * V (atoms per Å³; 0.15 by default, well below the
jamming density of 1.5 Å random sequential addition) is capped to that value.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from .forcefield import ForceField, Guest
from .structure import Framework, cell_matrix, perpendicular_lengths, replicate, volume
from .widom import KB

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
CHARGE_KEY = 0xC4A2_9E5D_1B07_F3A1

# Framework LJ types of the synthetic workloads: (symbol, epsilon/kB in K, sigma in Å).
FRAMEWORK_TYPES = (("C", 52.8, 3.43), ("H", 22.1, 2.57), ("O", 30.2, 3.12), ("Zn", 62.4, 2.46))


def _u64(x):
    return np.asarray(x, dtype=np.uint64)


def splitmix64_np(seed, index):
    """Vectorised src/state.jl:225-230 on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = _u64(seed) + _u64(index) * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def chain_words(chain_seed, cycle, move, nwords):
    """Words 0..nwords-1 of ChainRNG(chain_seed, cycle, move), broadcast over array arguments."""
    key = splitmix64_np(splitmix64_np(chain_seed, cycle), move)
    return np.stack([splitmix64_np(key, np.uint64(i)) for i in range(nwords)], axis=-1)


def unit_f64(raw):
    return (raw >> np.uint64(11)).astype(np.float64) * 1.1102230246251565e-16


def _mulhi64(a: int, b: int) -> int:
    return (a * b) >> 64


_synth_lib = None


def _positions(A, n, seed, min_sep, max_attempts):
    """Random sequential addition in C (csrc/synth.c), the rule in the module docstring."""
    global _synth_lib
    if _synth_lib is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpa_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python pureadsorb.jl_amd/build.py`")
        L = ctypes.CDLL(path)
        L.pa_synth_positions.restype = ctypes.c_int
        L.pa_synth_positions.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_void_p]
        _synth_lib = L
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.reshape(9))
    frac = np.zeros((max(n, 1), 3))
    st = _synth_lib.pa_synth_positions(Acm.ctypes.data, n, seed, float(min_sep), int(max_attempts), frac.ctypes.data)
    if st != n:
        raise RuntimeError(f"synthetic_framework: atom {-st - 1} not placed after {max_attempts} attempts")
    return frac[:n]


def synthetic_framework(a, b, c, alpha, beta, gamma, natoms, seed, *, min_sep=1.5, charge_sd=0.0,
                        types=("C", "H", "O", "Zn"), max_density=0.15, max_attempts=100_000) -> Framework:
    """One unit cell of `natoms` random atoms (see the module docstring for the exact rule)."""
    A = cell_matrix(a, b, c, alpha, beta, gamma)
    n = int(min(natoms, math.floor(max_density * volume(A))))
    seed = int(seed) & MASK64
    frac = _positions(A, n, seed, min_sep, max_attempts)
    tw = chain_words(np.uint64(seed), np.arange(1, n + 1, dtype=np.uint64), np.uint64(0), 1)[:, 0]
    tidx = [_mulhi64(int(w), len(types)) for w in tw]
    symbols = [types[t] for t in tidx]
    if charge_sd > 0 and n > 0:
        w = chain_words(np.uint64(seed ^ CHARGE_KEY), np.arange(1, n + 1, dtype=np.uint64), np.uint64(0), 2)
        u1, u2 = unit_f64(w[:, 0]), unit_f64(w[:, 1])
        z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)
        q = charge_sd * z
        q = q - q.mean()
    else:
        q = np.zeros(n)
    return Framework(A, frac, [f"{s}{i + 1}" for i, s in enumerate(symbols)], symbols, q)


def replication_for_width(fw: Framework, width: float) -> tuple[int, int, int]:
    """Smallest replication making every perpendicular width >= `width`."""
    L = perpendicular_lengths(fw.cell)
    return tuple(max(1, int(math.ceil(width / L[i] - 1e-12))) for i in range(3))


def synthetic_forcefield(cutoff: float, *, tail: bool = False, guest: str = "co2") -> ForceField:
    """Framework types C_/H_/O_/Zn_ plus the guest's types, epsilon given as epsilon/kB (K)."""
    names = [s + "_" for s, _, _ in FRAMEWORK_TYPES]
    eps = [e * KB for _, e, _ in FRAMEWORK_TYPES]
    sig = [s for _, _, s in FRAMEWORK_TYPES]
    if guest == "co2":
        names += ["C_co2", "O_co2"]
        eps += [27.0 * KB, 79.0 * KB]
        sig += [2.80, 3.05]
    elif guest == "ch4":
        names += ["CH4"]
        eps += [148.0 * KB]
        sig += [3.73]
    else:
        raise ValueError(f"unknown guest {guest!r}")
    return ForceField.from_types(names, sig, eps, cutoff=cutoff, tail=tail)


def co2_guest(ff: ForceField) -> Guest:
    """Rigid linear CO2: C (+0.70 e) at the centre, O (-0.35 e) at +-1.16 Å."""
    c, o = ff.names.index("C_co2"), ff.names.index("O_co2")
    return Guest([[0.0, 0.0, 0.0], [-1.16, 0.0, 0.0], [1.16, 0.0, 0.0]], [c, o, o], [0.70, -0.35, -0.35],
                 303.75, 7.84e6, 0.22394)


def ch4_guest(ff: ForceField) -> Guest:
    """United-atom CH4: one LJ site, no charge."""
    return Guest([[0.0, 0.0, 0.0]], [ff.names.index("CH4")], [0.0], 190.564, 4.5992e6, 0.01142)


def config0_framework(seed: int = 42) -> Framework:
    """Cubic a = 12.9 Å, 106 atoms, 1.5 Å separation, no charges, replicated 2x2x2 (848 atoms)."""
    fw = synthetic_framework(12.9, 12.9, 12.9, 90, 90, 90, 106, seed)
    return replicate(fw, (2, 2, 2))


def config1_framework(seed: int = 42) -> Framework:
    """Hexagonal 13/13/15 Å, gamma = 120, 300 atoms, Gaussian charges (sd 0.5 e) shifted to
    neutral, replicated 3x3x3 (8,100 atoms)."""
    fw = synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 300, seed, charge_sd=0.5)
    return replicate(fw, (3, 3, 3))


def config2_frameworks(n: int = 256, seed: int = 42, width: float = 24.0) -> list[Framework]:
    """`n` frameworks: natoms log-uniform in [200, 6000], lengths U(10, 40) Å, angles U(70, 110)
    degrees, charges as config 1, each replicated until every perpendicular width >= `width`."""
    out = []
    w = unit_f64(chain_words(np.uint64(seed ^ 0x5EED_F4A3), np.arange(1, n + 1, dtype=np.uint64), np.uint64(0), 7))
    for i in range(n):
        u = w[i]
        natoms = int(round(math.exp(math.log(200) + u[0] * (math.log(6000) - math.log(200)))))
        a, b, c = (10 + 30 * u[1], 10 + 30 * u[2], 10 + 30 * u[3])
        al, be, ga = (70 + 40 * u[4], 70 + 40 * u[5], 70 + 40 * u[6])
        try:
            fw = synthetic_framework(a, b, c, al, be, ga, natoms, seed * 1000003 + i + 1, charge_sd=0.5)
        except Exception:
            # a degenerate angle triple: redraw gamma deterministically toward 90 degrees
            fw = synthetic_framework(a, b, c, al, be, 90.0, natoms, seed * 1000003 + i + 1, charge_sd=0.5)
        out.append(replicate(fw, replication_for_width(fw, width)))
    return out
