"""The device move samplers against the distributions the mathematics fixes (the same
references as tests/test_oracle_sampling.py, with more chains): reinsertion and translation
against the Boltzmann radial density of a single-atom host (test/moves_tests.jl:333-414),
rotation against the orientation density of a rigid dumbbell at its fixed position, and
insertion / deletion against the Poisson loading of guests that only feel the host's field."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

import oracle as O
import pureadsorb_amd as pa
import sampling_refs as R

pytestmark = pytest.mark.gpu

EDGES = [0.0, 3.2, 3.6, 4.0, 4.5, 5.0]
CEDGES = [-1.0, -0.6, -0.2, 0.2, 0.6, 1.0]


def _state(nsys, sites, types, *, ncount=1, capacity=1, seed=7, guest_guest_eps=None):
    fw, ff, g, ew = R.single_atom_system(sites, types, guest_guest_eps=guest_guest_eps)
    b = pa.FrameworkBatch([fw] * nsys, ff, g, ew, fullk=True)
    st = pa.SystemState(b, g, [ncount] * nsys, T=R.T, seed=seed, capacities=[capacity] * nsys)
    return b, st


def _poses(st):
    ps, qs = [], []
    for n in range(st.nsys):
        p, q, _ = st.guests(n)
        ps.append(p)
        qs.append(q)
    return ps, qs


def _radii(ps):
    c = np.full(3, R.L_BOX / 2)
    return np.array([np.linalg.norm(R.minimum_image_cubic(x - c)) for p in ps for x in p])


@pytest.mark.parametrize("movetype,nmoves,step", [(pa.MOVE_REINSERTION, 40, 0.0), (pa.MOVE_TRANSLATION, 300, 1.5)])
def test_device_moves_sample_the_boltzmann_radial_density(movetype, nmoves, step):
    """16,000 chains, one guest each: after the moves the reference point's distance from the
    host atom follows exp(-U/kT) (chi-square p > 1e-4), and the same radii reject the uniform
    density that an always-accepting sampler (or one that never leaves its uniform start) gives."""
    nsys = 16_000
    print(f"device sampler: move type {movetype}, {nsys} chains x {nmoves}", flush=True)
    _, st = _state(nsys, [[0.0, 0.0, 0.0]], [1])
    pa.mc_moves(st, movetype, nmoves, T=R.T, step_trans=step)
    ps, _ = _poses(st)
    obs = R.categorize_radii(_radii(ps), EDGES)
    p_ref, _ = R.radial_probabilities(EDGES)
    pval, chi2 = R.chi2_pvalue(obs, p_ref)
    acc = sum(st.info(n)["accepted"][movetype - 1] for n in range(0, nsys, 97))
    att = sum(st.info(n)["attempted"][movetype - 1] for n in range(0, nsys, 97))
    print(f"observed {obs.tolist()}, expected {np.round(nsys * p_ref, 1).tolist()}, chi2 {chi2:.2f}, p {pval:.3g}, "
          f"acceptance {acc / att:.3f}", flush=True)
    assert pval > 1e-4
    assert R.chi2_pvalue(obs, R.uniform_radial_probabilities(EDGES))[0] < 1e-8
    assert 0.1 < acc / att < 0.99


@pytest.mark.parametrize("step,nmoves", [(1.0, 60), (0.4, 150)])
def test_device_rotation_samples_the_orientation_given_the_position(step, nmoves):
    """Rigid dumbbells (sites 1.5 Å apart), rotation moves only: every position stays where
    the state put it, and the orientation given that position follows exp(-U/kT). Pooled counts
    of u . r_hat over the chains whose reference site lies 3-6 Å from the host atom, against the
    per-chain quadrature probabilities (Mahalanobis chi-square)."""
    nsys = 12_000
    print(f"device rotation: step {step}, {nsys} chains x {nmoves}", flush=True)
    _, st = _state(nsys, [[0.0, 0.0, 0.0], [R.DUMBBELL_D, 0.0, 0.0]], [1, 1], seed=int(10 * step))
    p0, _ = _poses(st)
    pa.mc_moves(st, pa.MOVE_ROTATION, nmoves, T=R.T, step_rot=step)
    ps, qs = _poses(st)
    c = np.full(3, R.L_BOX / 2)
    plist, obs = [], np.zeros(len(CEDGES) - 1)
    for a, p, q in zip(p0, ps, qs):
        assert np.array_equal(a, p)
        d = R.minimum_image_cubic(p[0] - c)
        r = float(np.linalg.norm(d))
        if not 3.0 <= r <= 6.0:
            continue
        cos = float(np.dot(O.rotate(q[0], [1.0, 0.0, 0.0]), d / r))
        obs[min(np.searchsorted(CEDGES, cos, side="right") - 1, len(CEDGES) - 2)] += 1
        plist.append(R.orientation_probabilities(r, CEDGES))
    pval, stat = R.mahalanobis_pvalue(obs, plist)
    flat = R.mahalanobis_pvalue(obs, np.full((len(plist), len(CEDGES) - 1), 1.0 / (len(CEDGES) - 1)))[0]
    print(f"{len(plist)} chains: observed {obs.tolist()}, expected {np.round(np.sum(plist, axis=0), 1).tolist()}, "
          f"stat {stat:.2f}, p {pval:.3g}; uniform-orientation p {flat:.3g}", flush=True)
    assert len(plist) > 2000
    assert pval > 1e-4
    assert flat < 1e-8


def test_device_exchange_gives_a_poisson_ideal_gas_in_the_host_field():
    """GCMC with exchange moves only, guests with zero guest-guest epsilon: the loading of each
    chain is Poisson with mean f Z / kT (= 2 here). 4,000 independent chains after 80 cycles."""
    nsys = 4000
    _, Z = R.radial_probabilities(EDGES)
    mean_ref = 2.0
    f_pa = mean_ref * R.KT / Z / pa.PASCAL
    print(f"device exchange: {nsys} chains, fugacity {f_pa:.4g} Pa", flush=True)
    _, st = _state(nsys, [[0.0, 0.0, 0.0]], [1], ncount=0, capacity=40, seed=33, guest_guest_eps=0.0)
    res = pa.run_gcmc(st, T=R.T, n_warmup=60, n_production=20, n_audit=1000, step_trans=0.0, step_rot=0.0,
                      fugacity=f_pa, p_translation=0.0, p_rotation=0.0, p_reinsertion=0.0, p_exchange=1.0,
                      min_cycle_length=8, nblocks=2)
    ns = np.array([st.nguests(n) for n in range(nsys)])
    mean, var = ns.mean(), ns.var(ddof=1)
    pooled = float(np.mean([r.loading for r in res]))
    print(f"<N> {mean:.4f} (Poisson mean {mean_ref}), var {var:.4f}, production loading {pooled:.4f}", flush=True)
    assert abs(mean - mean_ref) < 5 * math.sqrt(mean_ref / nsys)
    assert abs(var - mean_ref) < 5 * math.sqrt((2 * mean_ref**2 + mean_ref) / nsys)
    assert abs(pooled - mean_ref) < 5 * math.sqrt(mean_ref / nsys)
    kmax = 8
    obs = np.array([np.sum(ns == k) for k in range(kmax)] + [np.sum(ns >= kmax)])
    pk = np.array([stats.poisson.pmf(k, mean_ref) for k in range(kmax)] + [stats.poisson.sf(kmax - 1, mean_ref)])
    assert R.chi2_pvalue(obs, pk)[0] > 1e-4
    assert abs(mean - mean_ref * R.L_BOX**3 / Z) > 10 * math.sqrt(mean_ref / nsys)
