"""Pins the oracle's Widom statistics (src/widom.jl:80-84, :290-492) against closed forms:
the empty box, one Lennard-Jones atom against the radial quadrature, the overflow guard and the
block remainder rule (test/widom_tests.jl:1-114,148-156)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O


def _single_site_system(L, atoms, sigma, eps, rc):
    A = np.diag([L, L, L])
    frac = np.array(atoms, dtype=float).reshape(-1, 3)
    n = len(frac)
    return O.OracleSystem(A, frac, [0] * n, [0.0] * n, (1, 1, 1), [[sigma]], [[eps]], rc, False,
                          [[0.0, 0.0, 0.0]], [0], [0.0], rc)


def test_boltzmann_weight_rules():
    """Underflow and an infinite energy give (0, 0), never NaN; a deep well stays finite in
    Float64 (test/widom_tests.jl:1-41)."""
    for dU, kT, ew, euw in ((1000.0, 0.0257, 0.0, 0.0), (np.inf, 0.0257, 0.0, 0.0),
                            (-100.0, 1.0, np.exp(100.0), -100 * np.exp(100.0))):
        w, uw = O.boltzmann_weight(dU, kT)
        assert w == ew and uw == euw and not np.isnan(uw)


def test_empty_box_ideal_gas():
    o = _single_site_system(30.0, np.zeros((0, 3)), 3.0, 0.001, 12.0)
    st, out, _ = o.widom(300.0, 10_000, 1, 0, 10)
    kT = O.KB * 300.0
    assert st == 0
    assert abs(out[0]) < 1e-12
    assert out[2] == pytest.approx(30.0**3 / kT, rel=1e-14)
    assert out[4] == pytest.approx(kT, rel=1e-14)
    assert out[6] == 10_000


def test_single_lj_atom_radial_integral():
    """<W> = 1 - 4 pi int_0^rc (1 - exp(-beta u)) r^2 dr / L^3, within 4 standard errors."""
    L, sig, eps, rc, T = 40.0, 3.4, 0.0103, 12.0, 300.0
    o = _single_site_system(L, [[0.5, 0.5, 0.5]], sig, eps, rc)
    st, out, _ = o.widom(T, 2_000_000, 2, 0, 20)
    assert st == 0
    beta = 1 / (O.KB * T)
    u = lambda x: 4 * eps * ((sig / x) ** 12 - (sig / x) ** 6)  # noqa: E731
    integral, _ = quad(lambda x: (1 - np.exp(-beta * u(x))) * x * x, 1e-3, rc, epsrel=1e-10, limit=200)
    expected = 1 - 4 * np.pi * integral / L**3
    meanW = out[2] * O.KB * T / L**3
    errW = out[3] * O.KB * T / L**3
    assert abs(meanW - expected) < 4 * errW


def test_reduce_overflow_and_empty_block():
    st, _ = O.reduce([np.inf, np.inf], [-np.inf, -np.inf], [50, 50], 1.0, 1.0)
    assert st == -2
    st, _ = O.reduce([1.0, 1.0], [0.5, 0.5], [50, 0], 1.0, 1.0)
    assert st == -1


def test_remainder_block_absorbs_extra_samples():
    o = _single_site_system(30.0, np.zeros((0, 3)), 3.0, 0.001, 12.0)
    st, out, bs = o.widom(300.0, 81, 4, 0, 10)
    assert st == 0 and out[6] == 81
    assert list(bs[:, 2]) == [8] * 9 + [9]


def test_rubtak_co2_against_recorded_reference(rubtak, trappe, co2):
    """The reference's recorded 2,000-insertion RUBTAK 3x3x3 + CO2 result (seed 3, 4 blocks,
    test/widom_tests.jl:274-298) came from a different random stream (Xoshiro), so agreement
    is statistical: within 3 combined standard errors."""
    import pureadsorb_amd as pa
    from conftest import oracle_for
    sc = pa.replicate(rubtak, (3, 3, 3))
    o = oracle_for(sc, trappe, co2, pa.EwaldParams(12.0, 1e-6))
    st, out, _ = o.widom(298.15, 2000, 3, 0, 4)
    assert st == 0
    ref = {"mu_ex": (-0.14435951741921793, 0.004098060610114027), "K_H": (6.590899117050657e8, 1.0512729413942294e8),
           "q_st": (0.2563139447257997, 0.0025536418864197745)}
    ours = {"mu_ex": (out[0], out[1]), "K_H": (out[2], out[3]), "q_st": (out[4], out[5])}
    for k, (m, s) in ref.items():
        assert abs(ours[k][0] - m) < 3 * np.hypot(ours[k][1], s), (k, ours[k], m, s)
