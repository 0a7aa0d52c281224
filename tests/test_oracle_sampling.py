"""The oracle's move samplers against distributions the mathematics fixes, so that a wrong
acceptance rule, proposal or random-variate transform in the oracle fails here rather than being
copied into the device trace it is compared with:

* rand_normal: mean, variance and fourth moment of the Box-Muller normal (test/rng_tests.jl:100-121);
* quaternion_pow / qmul / rotate: scipy's rotation-vector exponential and composition;
* reinsertion: the reference's chi-square test of a single-atom host (test/moves_tests.jl:333-414);
* translation: chains started uniform in the box converge to the same Boltzmann radial density;
* rotation: a rigid dumbbell's orientation, given its fixed position, follows exp(-U/kT);
* insertion / deletion: non-interacting guests in the host atom's field are a Poisson ideal gas
  with mean f Z / kT, Z = integral of exp(-U/kT) over the box.
Each sampler test also shows that the same data reject the distribution a plausible mistake
would produce (always accept: uniform; never accept: the start)."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats
from scipy.spatial.transform import Rotation

import oracle as O
import pureadsorb_amd as pa
import sampling_refs as R
from conftest import oracle_for

EDGES = [0.0, 3.2, 3.6, 4.0, 4.5, 5.0]


def test_rand_normal_moments():
    """test/rng_tests.jl:100-121 (5-sigma bands on the mean and variance) plus E[z^4] = 3
    (Var(z^4) = 105 - 9), which a Box-Muller radius off by a constant factor fails."""
    n = 2_000_000
    z = O.normals(321, 4, 4, n)
    assert abs(z.mean()) < 5 * math.sqrt(1 / n)
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / (n - 1))
    assert abs(np.mean(z ** 4) - 3) < 5 * math.sqrt(96 / n)
    # the whole distribution, on a subsample (Kolmogorov-Smirnov)
    assert stats.kstest(z[:200_000], "norm").pvalue > 1e-4


def test_rand_quaternion_is_haar_uniform_in_angle():
    """The rotation angle phi of a Haar-uniform rotation has CDF (phi - sin phi)/pi on [0, pi]."""
    q = O.quaternions(9, 5, 1, 200_000)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    phi = 2 * np.arccos(np.clip(np.abs(q[:, 3]), 0, 1))
    assert stats.kstest(phi, lambda x: (x - np.sin(x)) / np.pi).pvalue > 1e-4


def test_quaternion_pow_qmul_rotate_against_scipy():
    """quaternion_pow(q, t) is the rotation t*theta about q's axis, theta = 2 acos(w) (scipy's
    rotation-vector exponential); qmul(q2, q1) applies q1 first; rotate is the rotation of a
    vector by a unit quaternion in (x, y, z, w) order."""
    qs = O.quaternions(17, 1, 2, 300)
    rng = np.random.default_rng(0)
    for q in qs:
        v = rng.normal(size=3)
        assert np.allclose(O.rotate(q, v), Rotation.from_quat(q).apply(v), atol=1e-12)
        assert np.array_equal(O.quaternion_pow(q, 0.0), [0.0, 0.0, 0.0, 1.0])
        assert np.allclose(O.quaternion_pow(q, 1.0), q, atol=1e-12)
        th = 2 * math.acos(min(1.0, max(-1.0, q[3])))
        axis = q[:3] / np.linalg.norm(q[:3])
        for t in (0.1, 0.3, 0.5, 0.9):
            want = Rotation.from_rotvec(t * th * axis).as_matrix()
            got = Rotation.from_quat(O.quaternion_pow(q, t)).as_matrix()
            assert np.allclose(got, want, atol=1e-12)
        # one-parameter subgroup: q^a q^b = q^(a+b)
        assert np.allclose(Rotation.from_quat(O.qmul(O.quaternion_pow(q, 0.3), O.quaternion_pow(q, 0.5))).as_matrix(),
                           Rotation.from_quat(O.quaternion_pow(q, 0.8)).as_matrix(), atol=1e-12)
    for a, b in zip(qs[:100], qs[100:200]):
        want = (Rotation.from_quat(a) * Rotation.from_quat(b)).as_matrix()  # b first, then a
        assert np.allclose(Rotation.from_quat(O.qmul(a, b)).as_matrix(), want, atol=1e-12)
        v = rng.normal(size=3)
        assert np.allclose(O.rotate(O.qmul(a, b), v), O.rotate(a, O.rotate(b, v)), atol=1e-12)


def _single_site_oracle(guest_guest_eps=None):
    fw, ff, g, ew = R.single_atom_system([[0.0, 0.0, 0.0]], [1], guest_guest_eps=guest_guest_eps)
    return oracle_for(fw, ff, g, ew)


def _radii(chains):
    c = np.full(3, R.L_BOX / 2)
    rs = []
    for m in chains:
        p, _, _ = m.guests()
        for x in p:
            rs.append(np.linalg.norm(R.minimum_image_cubic(x - c)))
    return np.array(rs)


def test_reinsertion_samples_the_boltzmann_radial_density():
    """test/moves_tests.jl:333-414 on the oracle: 4,000 chains x 40 reinsertions from a corner
    start, chi-square p > 1e-4 against the quadrature reference; the same radii reject the
    uniform density of an always-accepting sampler."""
    o = _single_site_oracle()
    p_ref, _ = R.radial_probabilities(EDGES)
    chains = []
    for n in range(4000):
        m = O.OracleMC(o, 1, O.splitmix64(7, n + 1), np.zeros((1, 3)), np.array([[0.0, 0.0, 0.0, 1.0]]))
        for _ in range(40):
            m.attempt(3, T=R.T)
        chains.append(m)
    obs = R.categorize_radii(_radii(chains), EDGES)
    pval, chi2 = R.chi2_pvalue(obs, p_ref)
    print(f"reinsertion: observed {obs.tolist()}, expected {np.round(4000 * p_ref, 1).tolist()}, chi2 {chi2:.2f}, "
          f"p {pval:.3g}")
    assert pval > 1e-4
    assert R.chi2_pvalue(obs, R.uniform_radial_probabilities(EDGES))[0] < 1e-8


def test_translation_converges_to_the_boltzmann_radial_density():
    """Chains started uniform in the box, 300 Gaussian translations (s.d. 1.5 Å) each: the
    radial density is the Boltzmann one; a sampler that never accepts keeps the uniform start,
    which the same data reject."""
    o = _single_site_oracle()
    p_ref, _ = R.radial_probabilities(EDGES)
    rng = np.random.default_rng(11)
    chains = []
    for n in range(4000):
        x = rng.uniform(0.0, R.L_BOX, 3)
        if np.linalg.norm(x - R.L_BOX / 2) < 2.6:   # skip starts inside the core (energy ~ +inf)
            x = np.zeros(3)
        m = O.OracleMC(o, 1, O.splitmix64(21, n + 1), x[None, :], np.array([[0.0, 0.0, 0.0, 1.0]]))
        m.cycle((1.0, 0.0, 0.0, 0.0), 300, 1.5, 0.0, 1.0, R.T)
        chains.append(m)
    obs = R.categorize_radii(_radii(chains), EDGES)
    pval, chi2 = R.chi2_pvalue(obs, p_ref)
    acc, att = chains[0].counts()
    frac = sum(int(c.counts()[0][0]) for c in chains) / (300.0 * len(chains))
    print(f"translation: observed {obs.tolist()}, expected {np.round(4000 * p_ref, 1).tolist()}, chi2 {chi2:.2f}, "
          f"p {pval:.3g}, acceptance {frac:.3f}")
    assert pval > 1e-4
    assert 0.2 < frac < 0.99
    assert R.chi2_pvalue(obs, R.uniform_radial_probabilities(EDGES))[0] < 1e-8


CEDGES = [-1.0, -0.6, -0.2, 0.2, 0.6, 1.0]


def test_rotation_samples_the_orientation_given_the_position():
    """A rigid dumbbell (sites 1.5 Å apart, the reference site at the first) with rotation moves
    only: positions never change, and the orientation of each chain follows exp(-U/kT) given its
    position. Pooled counts of u . r_hat per bin against the per-chain quadrature probabilities
    (Mahalanobis chi-square over independent chains), for step 1 (uniform proposals) and 0.4."""
    fw, ff, g, ew = R.single_atom_system([[0.0, 0.0, 0.0], [R.DUMBBELL_D, 0.0, 0.0]], [1, 1])
    o = oracle_for(fw, ff, g, ew)
    c = np.full(3, R.L_BOX / 2)
    for step, nmoves, seed in ((1.0, 60, 5), (0.4, 150, 6)):
        rng = np.random.default_rng(seed)
        plist, obs, moved = [], np.zeros(len(CEDGES) - 1), 0
        while len(plist) < 1500:
            # reference sites at 3.0-6.0 Å from the host atom, where the second site's energy
            # depends on the orientation
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            x = c + rng.uniform(3.0, 6.0) * u
            q0 = O.quaternions(seed, len(plist) + 1, 9, 1)[0]
            m = O.OracleMC(o, 1, O.splitmix64(seed, len(plist) + 1), x[None, :], q0[None, :])
            m.cycle((0.0, 1.0, 0.0, 0.0), nmoves, 0.0, step, 1.0, R.T)
            p, q, _ = m.guests()
            d = R.minimum_image_cubic(p[0] - c)
            r = np.linalg.norm(d)
            assert np.allclose(p[0], x)
            axis = O.rotate(q[0], [1.0, 0.0, 0.0])
            cos = float(np.dot(axis, d / r))
            obs[min(np.searchsorted(CEDGES, cos, side="right") - 1, len(CEDGES) - 2)] += 1
            plist.append(R.orientation_probabilities(r, CEDGES))
            moved += int(m.counts()[0][1] > 0)
        pval, stat = R.mahalanobis_pvalue(obs, plist)
        flat = R.mahalanobis_pvalue(obs, np.full((len(plist), len(CEDGES) - 1), 1.0 / (len(CEDGES) - 1)))[0]
        print(f"rotation step {step}: observed {obs.tolist()}, expected {np.round(np.sum(plist, axis=0), 1).tolist()}, "
              f"stat {stat:.2f}, p {pval:.3g}; uniform-orientation p {flat:.3g}")
        assert pval > 1e-4
        assert flat < 1e-8
        assert moved > 0.9 * len(plist)


def test_exchange_gives_a_poisson_ideal_gas_in_the_host_field():
    """Guests with zero guest-guest epsilon only feel the host atom: the grand-canonical loading
    is Poisson with mean f Z / kT, Z = integral over the box of exp(-U/kT) (here ~40 % above the
    box volume, so the energy term of the acceptance matters as much as the prefactor)."""
    o = _single_site_oracle(guest_guest_eps=0.0)
    _, Z = R.radial_probabilities(EDGES)
    mean_ref = 2.0
    f_pa = mean_ref * R.KT / Z / pa.PASCAL
    ns = []
    for n in range(2000):
        m = O.OracleMC(o, 40, O.splitmix64(33, n + 1))
        for _ in range(60):
            m.cycle((0.0, 0.0, 0.0, 1.0), 8, 0.0, 0.0, f_pa, R.T)
        ns.append(m.scalars()["occupancy"])
    ns = np.array(ns)
    mean, var = ns.mean(), ns.var(ddof=1)
    print(f"exchange: <N> {mean:.4f} (Poisson mean {mean_ref}), var {var:.4f}, ideal gas without the field "
          f"{mean_ref * R.L_BOX**3 / Z:.4f}")
    assert abs(mean - mean_ref) < 5 * math.sqrt(mean_ref / len(ns))
    assert abs(var - mean_ref) < 5 * math.sqrt((2 * mean_ref**2 + mean_ref) / len(ns))
    kmax = 7
    obs = np.array([np.sum(ns == k) for k in range(kmax)] + [np.sum(ns >= kmax)])
    pk = np.array([stats.poisson.pmf(k, mean_ref) for k in range(kmax)] + [stats.poisson.sf(kmax - 1, mean_ref)])
    assert R.chi2_pvalue(obs, pk)[0] > 1e-4
    # the field-free ideal gas (energy term dropped) is rejected by the same data
    assert abs(mean - mean_ref * R.L_BOX**3 / Z) > 10 * math.sqrt(mean_ref / len(ns))
