"""Force field, guests and Ewald parameters (src/forcefield.jl, src/ewald.jl:8-18)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import yaml

from ._lib import ArgumentError, DimensionMismatch


@dataclass
class ForceField:
    """Per-pair Lennard-Jones tables (Å, eV) by Lorentz-Berthelot mixing (src/forcefield.jl:8-39)."""
    names: list
    sigma: np.ndarray
    epsilon: np.ndarray
    cutoff: float
    tail: bool = True

    @classmethod
    def from_types(cls, names, sigma, epsilon, *, cutoff, tail=True) -> "ForceField":
        s = np.asarray(sigma, dtype=np.float64)
        e = np.asarray(epsilon, dtype=np.float64)
        if not (len(names) == len(s) == len(e)):
            raise DimensionMismatch("names, sigma and epsilon must have the same length")
        S = (s[:, None] + s[None, :]) / 2
        E = np.sqrt(e[:, None] * e[None, :])
        return cls([str(n) for n in names], S, E, float(cutoff), bool(tail))

    @property
    def ntypes(self) -> int:
        return len(self.names)


def typeindex(ff: ForceField, name: str) -> int:
    """0-based index of an LJ type (src/forcefield.jl:41-45)."""
    try:
        return ff.names.index(name)
    except ValueError:
        raise ArgumentError(f"unknown LJ type {name}; force field defines {ff.names}") from None


def read_forcefield(path: str) -> ForceField:
    """kUPS-style YAML: `parameters: name: [sigma, epsilon]` (missing sigma = 1, epsilon = 0),
    `cutoff`, `tail_correction` (default true) (src/forcefield.jl:54-61)."""
    with open(path) as fh:
        d = yaml.safe_load(fh)
    params = d["parameters"]
    names = [str(n) for n in params.keys()]
    sig = [1.0 if params[n][0] is None else float(params[n][0]) for n in params]
    eps = [0.0 if params[n][1] is None else float(params[n][1]) for n in params]
    return ForceField.from_types(names, sig, eps, cutoff=float(d["cutoff"]), tail=bool(d.get("tail_correction", True)))


def _parse_number(x) -> float:
    return float(x.replace("_", "")) if isinstance(x, str) else float(x)


@dataclass
class Guest:
    """A rigid guest (src/forcefield.jl:76-83): site positions (N, 3) Å, LJ type indices
    (0-based), charges (e), critical temperature (K), pressure (Pa) and acentric factor."""
    sites: np.ndarray
    types: np.ndarray
    charges: np.ndarray
    tc: float = 1.0
    pc: float = 1.0
    omega: float = 0.0

    def __post_init__(self):
        self.sites = np.asarray(self.sites, dtype=np.float64).reshape(-1, 3)
        self.types = np.asarray(self.types, dtype=np.int32).reshape(-1)
        self.charges = np.asarray(self.charges, dtype=np.float64).reshape(-1)
        if not (len(self.sites) == len(self.types) == len(self.charges)):
            raise DimensionMismatch("guest sites, types and charges must have one entry per site")

    @property
    def nsites(self) -> int:
        return len(self.sites)


def read_guest(path: str, ff: ForceField) -> Guest:
    """kUPS-style guest YAML (src/forcefield.jl:91-101)."""
    with open(path) as fh:
        d = yaml.safe_load(fh)
    sites = np.array([[float(x) for x in p] for p in d["positions"]])
    types = [typeindex(ff, s) for s in d["symbols"]]
    return Guest(sites, types, [float(q) for q in d["charges"]], float(d["critical_temperature"]),
                 _parse_number(d["critical_pressure"]), float(d["acentric_factor"]))


def tail_coefficient(ff: ForceField, i: int, j: int) -> float:
    """src/forcefield.jl:103-109"""
    e = ff.epsilon[i, j]
    if e == 0:
        return 0.0
    s = ff.sigma[i, j]
    x = s / ff.cutoff
    return e * s**3 * (x**9 / 3 - x**3)


def tail_delta(ff: ForceField, counts, guest_counts, V) -> float:
    """Tail-correction change on adding one guest (src/forcefield.jl:111-124)."""
    if not ff.tail:
        return 0.0
    if len(counts) != ff.ntypes or len(guest_counts) != ff.ntypes:
        raise DimensionMismatch(f"counts must have one entry per LJ type ({ff.ntypes}), got {len(counts)}")
    acc = 0.0
    for a in range(ff.ntypes):
        for b in range(ff.ntypes):
            c = tail_coefficient(ff, a, b)
            acc += 2 * counts[a] * guest_counts[b] * c + guest_counts[a] * guest_counts[b] * c
    return 8 * math.pi / 3 / V * acc


@dataclass(frozen=True)
class EwaldParams:
    """Real-space cutoff (Å) and relative precision (src/ewald.jl:8-18)."""
    cutoff: float
    precision: float = 1.0e-6
