"""Frameworks, cells and CIF input (src/cell.jl, src/structure.jl)."""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import ArgumentError


def _sincosd(deg: float) -> tuple[float, float]:
    """sin/cos of an angle in degrees with exact zeros at multiples of 90 (Julia sind/cosd)."""
    n = round(deg / 90.0)
    r = math.radians(deg - 90.0 * n)
    s, c = math.sin(r), math.cos(r)
    q = n % 4
    if q == 0:
        return s, c
    if q == 1:
        return c, 0.0 - s
    if q == 2:
        return 0.0 - s, 0.0 - c
    return 0.0 - c, s


def cell_matrix(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """Cell matrix with lattice vectors as columns, a along x, b in the xy plane (src/cell.jl:2-12)."""
    if not (a > 0 and b > 0 and c > 0):
        raise ArgumentError(f"cell lengths must be positive, got a={a}, b={b}, c={c}")
    _, ca = _sincosd(alpha)
    _, cb = _sincosd(beta)
    sg, cg = _sincosd(gamma)
    if sg == 0:
        raise ArgumentError(f"cell angle γ={gamma} is degenerate: sin(γ) is zero")
    cy = (ca - cb * cg) / sg
    rad = 1.0 - cb * cb - cy * cy
    if not rad > 0:
        raise ArgumentError(f"cell angles α={alpha}, β={beta}, γ={gamma} describe a degenerate cell")
    cz = math.sqrt(rad)
    return np.array([[a, b * cg, c * cb], [0.0, b * sg, c * cy], [0.0, 0.0, c * cz]])


def volume(A) -> float:
    return abs(float(np.linalg.det(np.asarray(A, dtype=np.float64))))


def perpendicular_lengths(A) -> np.ndarray:
    """src/cell.jl:14-20"""
    A = np.asarray(A, dtype=np.float64)
    V = volume(A)
    a, b, c = A[:, 0], A[:, 1], A[:, 2]
    return np.array([V / np.linalg.norm(np.cross(b, c)), V / np.linalg.norm(np.cross(a, c)),
                     V / np.linalg.norm(np.cross(a, b))])


def min_multiplicity(A, cutoff) -> tuple[int, int, int]:
    """Smallest replication giving every perpendicular width >= 2*cutoff (src/cell.jl:24-27)."""
    L = perpendicular_lengths(A)
    return tuple(int(math.ceil(2 * cutoff / L[i])) for i in range(3))


def reciprocal_basis(A) -> np.ndarray:
    return 2 * math.pi * np.linalg.inv(np.asarray(A, dtype=np.float64)).T


@dataclass
class Framework:
    """A periodic host structure (src/structure.jl:14-24): cell (Å, lattice vectors as columns),
    fractional coordinates (natoms, 3), labels, element symbols, charges (e) and the
    replication already folded into cell/frac."""
    cell: np.ndarray
    frac: np.ndarray
    labels: list
    symbols: list
    charges: np.ndarray
    replication: tuple = (1, 1, 1)
    _key: tuple = field(default=None, repr=False, compare=False)

    def __post_init__(self):
        self.cell = np.asarray(self.cell, dtype=np.float64).reshape(3, 3)
        self.frac = np.asarray(self.frac, dtype=np.float64).reshape(-1, 3)
        self.charges = np.asarray(self.charges, dtype=np.float64).reshape(-1)
        self.labels = list(self.labels)
        self.symbols = list(self.symbols)
        self.replication = tuple(int(x) for x in self.replication)
        n = len(self.frac)
        if not (len(self.labels) == len(self.symbols) == len(self.charges) == n):
            raise ArgumentError("frac, labels, symbols and charges must have one entry per atom")

    @property
    def natoms(self) -> int:
        return len(self.frac)

    def total_charge(self) -> float:
        return float(np.sum(self.charges))

    def cartesian(self) -> np.ndarray:
        return self.frac @ self.cell.T


def natoms(fw: Framework) -> int:
    return fw.natoms


def total_charge(fw: Framework) -> float:
    return fw.total_charge()


def cartesian(fw: Framework) -> np.ndarray:
    return fw.cartesian()


def read_cif(path: str) -> Framework:
    """Minimal P1 CIF reader with required _atom_site_charge (src/structure.jl:37-69)."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh.read().splitlines()]

    def getval(key):
        for ln in lines:
            if ln.startswith(key + " ") or ln.startswith(key + "\t"):
                return ln[len(key):].strip()
        raise ArgumentError(f"CIF is missing {key}")

    sg = getval("_symmetry_space_group_name_H-M").strip("'\"")
    if sg.replace(" ", "") != "P1":
        raise ArgumentError(f"only P1 CIF files are supported, got space group {sg}")
    keys6 = ("_cell_length_a", "_cell_length_b", "_cell_length_c", "_cell_angle_alpha", "_cell_angle_beta",
             "_cell_angle_gamma")
    cell = cell_matrix(*(float(getval(k)) for k in keys6))
    try:
        hstart = lines.index("_atom_site_label")
    except ValueError:
        raise ArgumentError("CIF has no _atom_site_label loop") from None
    cols = []
    i = hstart
    while i < len(lines) and lines[i].startswith("_atom_site_"):
        cols.append(lines[i])
        i += 1
    if "_atom_site_charge" not in cols:
        raise ArgumentError("CIF has no _atom_site_charge column; partial charges are required")

    def col(name):
        return cols.index(name) if name in cols else None

    cx, cy, cz, cq = (col(n) for n in ("_atom_site_fract_x", "_atom_site_fract_y", "_atom_site_fract_z",
                                         "_atom_site_charge"))
    cl, cs = col("_atom_site_label"), col("_atom_site_type_symbol")
    if any(v is None for v in (cx, cy, cz, cl, cs)):
        raise ArgumentError("CIF atom_site loop lacks label, type_symbol or fractional coordinates")
    frac, labels, symbols, charges = [], [], [], []
    while i < len(lines) and lines[i] and not lines[i].startswith("_") and not lines[i].startswith("loop_"):
        f = lines[i].split()
        if len(f) != len(cols):
            raise ArgumentError(f"CIF atom row has {len(f)} fields, header has {len(cols)}")
        frac.append((float(f[cx]), float(f[cy]), float(f[cz])))
        labels.append(f[cl])
        symbols.append(f[cs])
        charges.append(float(f[cq]))
        i += 1
    return Framework(cell, np.array(frac).reshape(-1, 3), labels, symbols, np.array(charges))


def replicate(fw: Framework, n) -> Framework:
    """n[0] x n[1] x n[2] supercell; image (i, j, k) with i fastest, atoms innermost
    (src/structure.jl:79-92)."""
    n = tuple(int(x) for x in n)
    if not all(x >= 1 for x in n):
        raise ArgumentError(f"replication factors must be ≥ 1, got {n}")
    N = n[0] * n[1] * n[2]
    shifts = np.array([(i, j, k) for k in range(n[2]) for j in range(n[1]) for i in range(n[0])], dtype=np.float64)
    frac = ((fw.frac[None, :, :] + shifts[:, None, :]) / np.array(n, dtype=np.float64)).reshape(-1, 3)
    return Framework(fw.cell @ np.diag(np.array(n, dtype=np.float64)), frac, fw.labels * N, fw.symbols * N,
                     np.tile(fw.charges, N), tuple(a * b for a, b in zip(fw.replication, n)))
