"""Pins the oracle's insertion energies (src/energy.jl:13-82, src/reference.jl:31-67): the full
reference energy equals a difference of whole-system Ewald energies (test/energy_tests.jl:164-248),
the production contract (coupled k, averaged self term) stays within the self-term half-range of
it (test/reference_tests.jl:1-60), and an independent numpy brute force (explicit image search,
scipy erfc) reproduces the production energy."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.special import erfc

import oracle as O
import pureadsorb_amd as pa
from conftest import oracle_for


def _rand_pose(rng, A):
    q = rng.random(4) - 0.5
    return A @ rng.random(3), q / np.linalg.norm(q)


def _rotate(q, v):
    u, w = q[:3], q[3]
    return v + 2 * np.cross(u, np.cross(u, v) + w * v)


def test_rotate_identities():
    """test/energy_tests.jl:154-162: identity, norm preservation, 90 degrees about z maps x to y."""
    v = np.array([0.3, -1.2, 2.0])
    out = np.zeros(3)
    O.lib().or_rotate(np.array([0.0, 0.0, 0.0, 1.0]), v, out)
    assert np.array_equal(out, v)
    O.lib().or_rotate(np.array([0.0, 0.0, np.sqrt(0.5), np.sqrt(0.5)]), np.array([1.0, 0.0, 0.0]), out)
    assert np.allclose(out, [0.0, 1.0, 0.0], atol=1e-12)
    rng = np.random.default_rng(1)
    for _ in range(100):
        q = rng.random(4) - 0.5
        q /= np.linalg.norm(q)
        O.lib().or_rotate(q, v, out)
        assert abs(np.linalg.norm(out) - np.linalg.norm(v)) < 1e-14


def test_minimum_image_vs_image_shell():
    """Rounding in fractional coordinates finds the nearest of the 125 images in a +-2 shell
    for separations under half the smallest perpendicular width (test/cell_tests.jl:108-122)."""
    A = O.cell_matrix(14.7619, 14.80147, 14.76539, 59.84578, 60.04729, 59.8131) @ np.diag([3.0, 3.0, 3.0])
    Acm = np.ascontiguousarray(A.T.reshape(9))
    inv = np.zeros(9)
    O.lib().or_inv3(Acm, inv)
    rng = np.random.default_rng(2)
    half = pa.perpendicular_lengths(A).min() / 2
    shifts = np.array([(i, j, k) for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)]) @ A.T
    n = 0
    for _ in range(2000):
        d = A @ (rng.random(3) * 4 - 2)
        out = np.zeros(3)
        O.lib().or_minimum_image(Acm, inv, d, out)
        best = np.min(np.linalg.norm(d + shifts, axis=1))
        if best < half:
            n += 1
            assert np.linalg.norm(out) == pytest.approx(best, rel=1e-12)
    assert n > 100


def _ff_rubtak():
    return pa.ForceField.from_types(["Zr_", "H_", "C_", "O_", "C", "O"], [2.78, 2.57, 3.43, 3.12, 2.8, 3.05],
                                    [0.003, 0.0019, 0.0046, 0.0026, 0.0023, 0.0068], cutoff=10.0, tail=False)


@pytest.mark.parametrize("rc_lj,rc_ew,rep", [(10.0, 10.0, None), (6.0, 12.0, (3, 3, 3))])
def test_reference_energy_is_full_system_difference(rubtak, co2, rc_lj, rc_ew, rep):
    """insertion_energy_reference = LJ + [E_ewald(host+guest) - E_ewald(host)] - guest self,
    exclusion and net-charge terms, rtol 1e-8 (test/energy_tests.jl:164-248)."""
    ff = _ff_rubtak()
    ff = pa.ForceField(ff.names, ff.sigma, ff.epsilon, rc_lj, False)
    g = pa.Guest(co2.sites, [4, 5, 5], co2.charges)
    sc = pa.replicate(rubtak, rep or pa.min_multiplicity(rubtak.cell, 10.0))
    o = oracle_for(sc, ff, g, pa.EwaldParams(rc_ew, 1e-7))
    A, V = sc.cell, o.V
    hpos, hq = sc.cartesian(), sc.charges
    htype = [ff.names.index(s + "_") for s in sc.symbols]
    alpha = o.alpha
    ks, w, _ = O.kvectors(A, O.ewald_kmax(alpha, 1e-7))
    rng = np.random.default_rng(7)
    pos, q = _rand_pose(rng, A)
    dU = o.insertion_energy_reference(pos, q)
    gpos = np.array([pos + _rotate(q, s) for s in g.sites])
    mol = np.concatenate([np.arange(1, len(hpos) + 1), np.zeros(3, dtype=int)])
    Ec = (O.ewald_energy(A, np.vstack([hpos, gpos]), np.concatenate([hq, g.charges]), mol, alpha, rc_ew, ks, w)
          - O.ewald_energy(A, hpos, hq, mol[: len(hpos)], alpha, rc_ew, ks, w))
    invA = np.linalg.inv(A)
    Elj = 0.0
    for sp, t in zip(gpos, g.types):
        d = sp - hpos
        d = d - (np.rint(d @ invA.T)) @ A.T
        r = np.linalg.norm(d, axis=1)
        m = r < rc_lj
        sig = ff.sigma[t, htype][m]
        eps = ff.epsilon[t, htype][m]
        x = (sig / r[m]) ** 6
        Elj += np.sum(4 * eps * (x * x - x))
    Eself = -O.KE * alpha / np.sqrt(np.pi) * np.sum(g.charges**2)
    Eexcl = 0.0
    for a in range(3):
        for b in range(a + 1, 3):
            r = np.linalg.norm(gpos[a] - gpos[b])
            Eexcl -= O.KE * g.charges[a] * g.charges[b] * (1 - O.erfc_dev(alpha * r)) / r
    Enet = -O.KE * np.pi / (2 * V * alpha**2) * ((hq.sum() + g.charges.sum()) ** 2 - hq.sum() ** 2)
    assert dU == pytest.approx(Elj + Ec - Eself - Eexcl - Enet, rel=1e-8)


def test_production_energy_within_self_term_halfrange(rubtak, trappe, co2):
    """Coupled-k cross term + orientation-averaged self term (the FrameworkBatch contract) vs the
    full reference sum, within 2*self_term_halfrange + 1e-12 relative, for CO2 and an SPC-like
    polar guest (test/reference_tests.jl:1-60)."""
    polar = pa.Guest([[0.0, 0.0, 0.0], [0.8165, 0.5774, 0.0], [-0.8165, 0.5774, 0.0]], co2.types, [-0.82, 0.41, 0.41])
    sc = pa.replicate(rubtak, (3, 3, 3))
    rng = np.random.default_rng(21)
    for g in (co2, polar):
        o = oracle_for(sc, trappe, g, pa.EwaldParams(12.0, 1e-6))
        old_offset = o.tail_term + O.KE * (o.self_term + o.excl_term + o.net_term)
        for _ in range(40):
            pos, q = _rand_pose(rng, sc.cell)
            prod = o.insertion_energy(pos, q)[0] + o.constant_offset
            ref = o.insertion_energy_reference(pos, q) + old_offset
            assert abs(prod - ref) <= 2 * o.self_halfrange + 1e-12 * abs(ref)


def test_production_energy_vs_numpy_brute_force(rubtak, trappe, co2):
    """An independent restatement: every image in a +-1 shell (nearest kept), scipy erfc,
    explicit cross term over the coupled k table; agreement to 1e-10 relative."""
    sc = pa.replicate(rubtak, (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    o = oracle_for(sc, trappe, co2, ew)
    A = sc.cell
    hpos, hq = sc.cartesian(), sc.charges
    htype = np.array([trappe.names.index(s + "_") for s in sc.symbols])
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]) @ A.T
    ks, kp, S, _ = o.ktable()
    rng = np.random.default_rng(9)
    checked = 0
    for _ in range(60):
        pos, q = _rand_pose(rng, A)
        gpos = np.array([pos + _rotate(q, s) for s in co2.sites])
        E = 0.0
        for s in range(3):
            d = (gpos[s] - hpos)[:, None, :] + shifts[None, :, :]
            r = np.min(np.linalg.norm(d, axis=2), axis=1)
            m = r < 12.0
            sig = trappe.sigma[co2.types[s], htype[m]]
            eps = trappe.epsilon[co2.types[s], htype[m]]
            x = (sig / r[m]) ** 6
            E += np.sum(4 * eps * (x * x - x))
            E += O.KE * co2.charges[s] * np.sum(hq[m] * erfc(o.alpha * r[m]) / r[m])
        Sg = (co2.charges[None, :] * np.exp(1j * ks @ gpos.T)).sum(axis=1)
        E += O.KE * np.sum(kp * 2 * np.real(np.conj(S) * Sg))
        e, npair = o.insertion_energy(pos, q)
        if abs(E) < 1e6:
            checked += 1
            assert e == pytest.approx(E, rel=1e-10, abs=1e-12)
    assert checked > 20
