"""Pins the oracle's Ewald machinery (src/ewald.jl:21-203): the NaCl Madelung constant,
independence of the splitting parameter, brute-force k-vector enumeration and the reference's
RUBTAK k-vector counts (test/ewald_tests.jl, test/batch_tests.jl:71-87)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.special import erfc

import oracle as O
import pureadsorb_amd as pa
from conftest import oracle_for


def test_alpha_and_kmax_follow_kups():
    a = O.ewald_alpha(12.0, 1.0e-6)
    assert erfc(a * 12.0) == pytest.approx(12.0 * 5.0e-7, rel=1e-9)
    assert O.ewald_kmax(a, 1.0e-6) == pytest.approx(2 * a * np.sqrt(-np.log(5.0e-7)), rel=1e-15)


@pytest.mark.parametrize("frac_rc", [0.49, 2.0])
def test_madelung_constant_of_nacl(frac_rc):
    """E per ion pair = -1.747565 KE / (a/2) (test/ewald_tests.jl:31-53)."""
    a = 5.64
    A = np.diag([a, a, a])
    base = [((0, 0, 0), 1.0), ((0.5, 0.5, 0), 1.0), ((0.5, 0, 0.5), 1.0), ((0, 0.5, 0.5), 1.0),
            ((0.5, 0, 0), -1.0), ((0, 0.5, 0), -1.0), ((0, 0, 0.5), -1.0), ((0.5, 0.5, 0.5), -1.0)]
    rc, prec = a * frac_rc, 1e-8
    m = pa.min_multiplicity(A, rc)
    Asc = A @ np.diag(m)
    pos, q = [], []
    for k in range(m[2]):
        for j in range(m[1]):
            for i in range(m[0]):
                for f, c in base:
                    pos.append(A @ (np.array(f) + [i, j, k]))
                    q.append(c)
    alpha = O.ewald_alpha(rc, prec)
    ks, w, _ = O.kvectors(Asc, O.ewald_kmax(alpha, prec))
    E = O.ewald_energy(Asc, np.array(pos), q, np.arange(len(q)), alpha, rc, ks, w)
    assert E / (4 * np.prod(m)) == pytest.approx(-1.747565 * O.KE / (a / 2), rel=1e-5)


def test_energy_independent_of_alpha():
    """test/ewald_tests.jl:55-68"""
    A = np.diag([20.0, 20.0, 20.0])
    rng = np.random.default_rng(3)
    pos = rng.random((40, 3)) @ A.T
    q = np.array([0.5 if i % 2 == 0 else -0.5 for i in range(40)])
    Es = []
    for rc in (6.0, 8.0, 9.5):
        alpha = O.ewald_alpha(rc, 1e-8)
        ks, w, _ = O.kvectors(A, O.ewald_kmax(alpha, 1e-8))
        Es.append(O.ewald_energy(A, pos, q, np.arange(40), alpha, rc, ks, w))
    assert all(e == pytest.approx(Es[0], rel=1e-6) for e in Es)


def test_intramolecular_exclusion_removes_the_pair():
    """test/ewald_tests.jl:112-122"""
    A = np.diag([30.0, 30.0, 30.0])
    pos = np.array([[15.0, 15.0, 15.0], [16.16, 15.0, 15.0]])
    q = [0.7, -0.35]
    alpha = O.ewald_alpha(12.0, 1e-8)
    ks, w, _ = O.kvectors(A, O.ewald_kmax(alpha, 1e-8))
    same = O.ewald_energy(A, pos, q, [1, 1], alpha, 12.0, ks, w)
    diff = O.ewald_energy(A, pos, q, [1, 2], alpha, 12.0, ks, w)
    assert diff - same == pytest.approx(O.KE * 0.7 * -0.35 / 1.16, rel=1e-3)


def test_kvectors_match_brute_force_on_triclinic_cell():
    """Enumeration bound from lattice-vector lengths, not perpendicular widths
    (test/ewald_tests.jl:70-98); integer coefficients reproduce the vectors (:100-110)."""
    A = O.cell_matrix(14.7619, 14.80147, 14.76539, 59.84578, 60.04729, 59.8131)
    alpha = O.ewald_alpha(12.0, 1e-6)
    kmax = O.ewald_kmax(alpha, 1e-6)
    B = 2 * np.pi * np.linalg.inv(A).T
    lat = np.linalg.norm(A, axis=0)
    perp = pa.perpendicular_lengths(A)
    n_lat = np.ceil(kmax * lat / (2 * np.pi)).astype(int)
    n_perp = np.ceil(kmax * perp / (2 * np.pi)).astype(int)
    assert np.all(n_lat >= n_perp) and np.any(n_lat > n_perp)
    N = int(n_lat.max())
    rng = np.arange(-2 * N, 2 * N + 1)
    n = np.array(np.meshgrid(np.arange(0, 2 * N + 1), rng, rng, indexing="ij")).reshape(3, -1).T
    n = n[~np.all(n == 0, axis=1)]
    k = n @ B.T
    keep = np.linalg.norm(k, axis=1) <= kmax
    ks, w, co = O.kvectors(A, kmax)
    assert len(ks) == keep.sum()
    assert w.sum() == np.where(n[keep, 0] == 0, 1.0, 2.0).sum()
    assert np.allclose(co @ B.T, ks, rtol=0, atol=1e-12)


def test_rubtak_3x3x3_coupled_k_counts(rubtak, trappe, co2):
    """4,587 k-vectors for the RUBTAK 3x3x3 cell, 190 coupled to the replication, and the host
    structure factor of every uncoupled one below 1e-9 (test/batch_tests.jl:71-87)."""
    sc = pa.replicate(rubtak, (3, 3, 3))
    o = oracle_for(sc, trappe, co2, pa.EwaldParams(12.0, 1e-6))
    assert (o.nkfull, o.nk) == (4587, 190)
    ks, kp, S, co = o.ktable(full=True)
    coupled = np.all(co % np.array(sc.replication) == 0, axis=1)
    assert coupled.sum() == 190
    assert np.all(np.abs(S[~coupled]) < 1e-9)
    kc, kpc, Sc, _ = o.ktable(full=False)
    assert np.array_equal(kc, ks[coupled]) and np.array_equal(Sc, S[coupled])
    # kprefactor = w * pk(|k|^2, alpha, V) with w = 1 for n1 == 0 else 2 (src/batch.jl; test/batch_tests.jl:22-29)
    w = np.where(co[:, 0] == 0, 1.0, 2.0)
    k2 = np.sum(ks * ks, axis=1)
    pk = 2 * np.pi / o.V * np.exp(-k2 / (4 * o.alpha**2)) / k2
    assert np.allclose(kp, w * pk, rtol=1e-14, atol=0)


def test_structure_factor_direct_sum(rubtak):
    """S(k) = sum_j q_j exp(i k.r_j) (src/ewald.jl:159-170) against numpy on the unit cell."""
    pos = rubtak.cartesian()
    alpha = O.ewald_alpha(12.0, 1e-6)
    ks, w, _ = O.kvectors(rubtak.cell, O.ewald_kmax(alpha, 1e-6))
    ks = ks[:200]
    Sre, Sim = np.zeros(len(ks)), np.zeros(len(ks))
    O.lib().or_structure_factor(len(ks), np.ascontiguousarray(ks.reshape(-1)), len(pos),
                                np.ascontiguousarray(pos.reshape(-1)), np.ascontiguousarray(rubtak.charges), Sre, Sim)
    ref = (rubtak.charges[None, :] * np.exp(1j * ks @ pos.T)).sum(axis=1)
    assert np.allclose(Sre + 1j * Sim, ref, rtol=0, atol=1e-12)
