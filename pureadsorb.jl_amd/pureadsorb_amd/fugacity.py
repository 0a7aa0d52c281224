"""Peng-Robinson fugacity of a pure gas (src/fugacity.jl), the host-side pressure -> fugacity
conversion of the μVT drivers. Runs once per system per run on the host, as in the reference."""
from __future__ import annotations

import math
from dataclasses import dataclass

from ._lib import ArgumentError


def cubic_roots(c2: float, c1: float, c0: float) -> list[complex]:
    """The three roots of Z^3 + c2 Z^2 + c1 Z + c0 (src/fugacity.jl:16-38): with Z = t - c2/3 the
    depressed cubic t^3 + p t + q has discriminant D = (q/2)^2 + (p/3)^3; D >= 0 gives one real
    root by Cardano's formula and a complex-conjugate pair, D < 0 three real roots by the
    trigonometric form."""
    p = c1 - c2 * c2 / 3.0
    q = 2.0 * c2 ** 3 / 27.0 - c2 * c1 / 3.0 + c0
    shift = c2 / 3.0
    disc = (q / 2.0) ** 2 + (p / 3.0) ** 3
    if disc >= 0.0:
        s = math.sqrt(disc)
        u = math.copysign(abs(-q / 2.0 + s) ** (1.0 / 3.0), -q / 2.0 + s)
        v = math.copysign(abs(-q / 2.0 - s) ** (1.0 / 3.0), -q / 2.0 - s)
        re = -(u + v) / 2.0 - shift
        im = (u - v) * math.sqrt(3.0) / 2.0
        return [complex(u + v - shift, 0.0), complex(re, im), complex(re, -im)]
    r = math.sqrt(-p ** 3 / 27.0)
    theta = math.acos(min(1.0, max(-1.0, -q / (2.0 * r))))
    m = 2.0 * math.sqrt(-p / 3.0)
    return [complex(m * math.cos((theta + 2.0 * k * math.pi) / 3.0) - shift, 0.0) for k in range(3)]


def log_fugacity_coefficient(Z: float, A: float, B: float) -> float:
    """Pure-component Peng-Robinson ln phi at compressibility Z (src/fugacity.jl:44-45)."""
    r2 = math.sqrt(2.0)
    return (Z - 1.0) - math.log(Z - B) - A / (2.0 * r2 * B) * math.log((Z + (1.0 + r2) * B) / (Z + (1.0 - r2) * B))


@dataclass(frozen=True)
class Fugacity:
    f: float
    phi: float
    Z: float


def peng_robinson_fugacity(P, Tgas, tc=None, pc=None, omega=None, *, guest=None) -> Fugacity:
    """`peng_robinson_fugacity(P, Tgas, tc, pc, omega)` or `(P, Tgas, guest)` (src/fugacity.jl:82-122):
    fugacity f (units of P), coefficient phi = f/P and compressibility Z. Of the real roots with
    Z > B it keeps the one of lowest ln phi (the stable phase)."""
    if guest is not None or (tc is not None and pc is None and omega is None and hasattr(tc, "tc")):
        g = guest if guest is not None else tc
        tc, pc, omega = g.tc, g.pc, g.omega
    P, Tgas, tc, pc, omega = float(P), float(Tgas), float(tc), float(pc), float(omega)
    if not P > 0.0:
        raise ArgumentError(f"peng_robinson_fugacity: P={P} must be positive")
    if not Tgas > 0.0:
        raise ArgumentError(f"peng_robinson_fugacity: Tgas={Tgas} must be positive")
    if not tc > 0.0:
        raise ArgumentError(f"peng_robinson_fugacity: tc={tc} must be positive")
    if not pc > 0.0:
        raise ArgumentError(f"peng_robinson_fugacity: pc={pc} must be positive")
    Pr, Tr = P / pc, Tgas / tc
    m = 0.37464 + 1.54226 * omega - 0.26992 * omega * omega
    sa = 1.0 + m * (1.0 - math.sqrt(Tr))
    A = 0.45724 * Pr / (Tr * Tr) * sa * sa
    B = 0.0778 * Pr / Tr
    roots = cubic_roots(B - 1.0, A - 3.0 * B * B - 2.0 * B, -(A * B - B * B - B ** 3))
    eps = 2.220446049250313e-16
    tol_im, tol_re = math.sqrt(eps), eps ** 0.75
    best_z, best = float("nan"), float("inf")
    for r in roots:
        z = r.real
        if not (abs(r.imag) < tol_im and z > B + tol_re):
            continue
        lp = log_fugacity_coefficient(z, A, B)
        if lp < best:
            best, best_z = lp, z
    if not math.isfinite(best):
        raise ArgumentError(f"peng_robinson_fugacity: no root has Z > B (an unphysical state point) at P={P}, "
                            f"Tgas={Tgas}, tc={tc}, pc={pc}, omega={omega}")
    phi = math.exp(best)
    return Fugacity(phi * P, phi, best_z)
