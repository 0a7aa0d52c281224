"""Closed-form and quadrature references for the move samplers' distribution tests, shared by the
oracle tests (CPU) and the device tests.

The host is a single Lennard-Jones atom at the centre of a cubic box of edge L, with the LJ
cutoff rc below L/2: the ball of radius rc around the atom never reaches a face, so the shell
volume at distance r < rc is exactly 4 pi r^2 and every point farther than rc has U = 0
(the reference's own construction, test/moves_tests.jl:333-414)."""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate, special

import pureadsorb_amd as pa

L_BOX = 14.0
SIGMA = 3.0
EPS = 0.08      # eV
RC = 5.0        # Å, < L/2
T = 298.15
KT = pa.KB * T


def u_lj(r, sigma=SIGMA, eps=EPS, rc=RC):
    """Truncated LJ of one site pair (zero at and beyond the cutoff, src/energy.jl:20)."""
    r = np.asarray(r, dtype=np.float64)
    x = (sigma / np.maximum(r, 1e-300)) ** 6
    return np.where(r < rc, 4.0 * eps * (x * x - x), 0.0)


def radial_weight(r):
    return 4.0 * math.pi * r * r * math.exp(-float(u_lj(r)) / KT)


def radial_probabilities(edges):
    """Probabilities of the reference point's distance to the host atom falling in each
    [edges[i], edges[i+1]) (edges[-1] == RC) plus the outside-cutoff category, under
    exp(-U/kT) over the box."""
    zin = integrate.quad(radial_weight, 0.0, RC, epsabs=0.0, epsrel=1e-11, limit=400, points=[SIGMA])[0]
    vout = L_BOX ** 3 - 4.0 / 3.0 * math.pi * RC ** 3
    Z = zin + vout
    p = [integrate.quad(radial_weight, edges[i], edges[i + 1], epsabs=0.0, epsrel=1e-11, limit=400)[0] / Z
         for i in range(len(edges) - 1)]
    return np.array(p + [vout / Z]), Z


def uniform_radial_probabilities(edges):
    """The same categories for a point uniform in the box (what an always-accepting sampler, or
    one that never moves from a uniform start, would give)."""
    V = L_BOX ** 3
    p = [4.0 / 3.0 * math.pi * (edges[i + 1] ** 3 - edges[i] ** 3) / V for i in range(len(edges) - 1)]
    return np.array(p + [1.0 - 4.0 / 3.0 * math.pi * RC ** 3 / V])


def categorize_radii(rs, edges):
    """Counts per category of radial_probabilities."""
    rs = np.asarray(rs)
    obs = np.zeros(len(edges), dtype=np.int64)
    for r in rs:
        if r >= edges[-1]:
            obs[-1] += 1
        else:
            obs[min(max(np.searchsorted(edges, r, side="right") - 1, 0), len(edges) - 2)] += 1
    return obs


def chi2_pvalue(obs, p):
    """Pearson chi-square goodness of fit of counts `obs` to probabilities `p` (upper tail)."""
    obs = np.asarray(obs, dtype=np.float64)
    n = obs.sum()
    e = n * np.asarray(p)
    chi2 = float(np.sum((obs - e) ** 2 / e))
    dof = len(obs) - 1
    return float(special.gammaincc(dof / 2.0, chi2 / 2.0)), chi2


def minimum_image_cubic(d, L=L_BOX):
    d = np.asarray(d, dtype=np.float64)
    return d - L * np.rint(d / L)


# ---- dumbbell guest: orientation given the reference point's distance r --------------------
DUMBBELL_D = 1.5  # Å between the two sites


def _u2(r, c, d=DUMBBELL_D):
    """LJ of the second site at p + d u (u . r_hat = c) with the host atom."""
    r2 = np.sqrt(np.maximum(r * r + d * d + 2.0 * r * d * c, 0.0))
    return u_lj(r2)


def orientation_probabilities(r, cedges, d=DUMBBELL_D):
    """P(u . r_hat in [cedges[i], cedges[i+1])) for a rigid dumbbell whose reference site sits at
    distance r from the host atom, orientation uniform on the sphere weighted by
    exp(-U_second_site/kT). The reference site's own energy does not depend on orientation."""
    # cos(theta) is uniform on [-1, 1] for a uniform direction. The integrand is smooth except
    # where the second site crosses the cutoff (the truncated LJ jumps there), so each bin is
    # split at that point and integrated by 48-point Gauss-Legendre on 16 panels; weights are
    # shifted by their minimum so that deep overlaps never underflow the normalisation.
    cstar = (RC * RC - r * r - d * d) / (2.0 * r * d)
    xg, wg = np.polynomial.legendre.leggauss(48)
    parts = []
    for i in range(len(cedges) - 1):
        a, b = cedges[i], cedges[i + 1]
        cuts = [a] + ([cstar] if a < cstar < b else []) + [b]
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            for p0, p1 in zip(np.linspace(lo, hi, 17)[:-1], np.linspace(lo, hi, 17)[1:]):
                parts.append((i, 0.5 * (p1 - p0) * xg + 0.5 * (p1 + p0), 0.5 * (p1 - p0) * wg))
    u = [_u2(r, c, d) for _, c, _ in parts]
    umin = min(float(np.min(x)) for x in u)
    acc = np.zeros(len(cedges) - 1)
    for (i, _, w), ux in zip(parts, u):
        acc[i] += float(np.sum(w * np.exp(-(ux - umin) / KT)))
    return acc / acc.sum()


def mahalanobis_pvalue(obs, plist):
    """Goodness of fit of pooled category counts of independent trials with per-trial
    probabilities plist[i] (rows sum to 1): (O - E)^T Sigma^-1 (O - E) over all categories but
    the last, Sigma = sum_i diag(p_i) - p_i p_i^T, which is asymptotically chi-square with
    (categories - 1) degrees of freedom."""
    P = np.asarray(plist, dtype=np.float64)
    obs = np.asarray(obs, dtype=np.float64)
    E = P.sum(axis=0)
    S = np.diag(E) - P.T @ P
    k = len(E) - 1
    d = (obs - E)[:k]
    stat = float(d @ np.linalg.solve(S[:k, :k], d))
    return float(special.gammaincc(k / 2.0, stat / 2.0)), stat


def single_atom_system(guest_sites, guest_types, nfw=1, guest_guest_eps=None):
    """Framework (one atom, type 0, at the box centre), force field and guest. Types: 0 host
    atom, 1 guest site. With guest_guest_eps set, the guest-guest epsilon is that value instead
    of the Lorentz-Berthelot one (0 makes guests an ideal gas in the host's field)."""
    cell = np.diag([L_BOX] * 3)
    fw = pa.Framework(cell, np.array([[0.5, 0.5, 0.5]]), ["X1"], ["X"], np.array([0.0]))
    S = np.full((2, 2), SIGMA)
    E = np.full((2, 2), EPS)
    if guest_guest_eps is not None:
        E[1, 1] = guest_guest_eps
    ff = pa.ForceField(["X_", "G_"], S, E, RC, False)
    g = pa.Guest(np.asarray(guest_sites, dtype=np.float64), list(guest_types), [0.0] * len(guest_types))
    return fw, ff, g, pa.EwaldParams(RC, 1e-6)
