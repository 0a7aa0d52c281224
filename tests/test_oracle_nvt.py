"""Pins the oracle's NVT pieces (src/nvt.jl): insertion_constant_term, the test insertion into a
loaded configuration (widom_chain_kernel!) and the run_nvt! restatement.

The insertion energy is pinned against an independent route: the difference of two from-scratch
total energies (the configuration with and without the test guest, structure factors rebuilt
from the poses), corrected by the exact per-orientation reciprocal self term that the insertion
replaces with its 64-orientation mean (the reference's own nvt_tests.jl oracle,
test/nvt_tests.jl:32-146, compares the same two quantities on average). The driver is pinned by
the reference's own checks: with no guests it reproduces Widom on the same insertion stream
(test/nvt_tests.jl:182-210), the cycle length is max(max occupancy, min_cycle_length) shared by
the batch (:212-240)."""
import math

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from pureadsorb_amd import synth

from conftest import oracle_for

T = 298.15


def _system(tail=True, charged=True, natoms=30):
    ff = synth.synthetic_forcefield(11.0, tail=tail)
    g = synth.co2_guest(ff)
    if not charged:
        g = pa.Guest(g.sites, g.types, [0.0, 0.0, 0.0])
    uc = synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, natoms, 7, charge_sd=0.5)
    fw = pa.replicate(uc, (2, 2, 2))
    o = oracle_for(fw, ff, g, pa.EwaldParams(11.0))
    o.guest = g
    return o


def _poses(o, n, seed, minsep=3.5, host_negative=True):
    rng = np.random.default_rng(seed)
    pos, quat = [], []
    while len(pos) < n:
        p = o.cell @ rng.random(3)
        u = rng.random(3)
        q = np.array([math.sqrt(1 - u[0]) * math.cos(2 * math.pi * u[1]), math.sqrt(u[0]) * math.sin(2 * math.pi * u[2]),
                      math.sqrt(u[0]) * math.cos(2 * math.pi * u[2]), math.sqrt(1 - u[0]) * math.sin(2 * math.pi * u[1])])
        if host_negative and O.lib().or_host_guest(o._h, np.ascontiguousarray(p), np.ascontiguousarray(q)) >= 0.0:
            continue
        if any(np.linalg.norm(p - x) < minsep for x in pos):
            continue
        pos.append(p)
        quat.append(q)
    return np.array(pos), np.array(quat)


def _rotate(q, v):
    u, w = np.asarray(q[:3]), q[3]
    t = np.cross(u, v) + w * np.asarray(v)
    return np.asarray(v) + 2.0 * np.cross(u, t)


def _exact_self(o, q):
    """KE sum_k kpref |S_g(k)|^2 of one orientation over the full k table (numpy)."""
    ks, kp, _, _ = o.ktable(full=True)
    sites = np.array([_rotate(q, s) for s in np.asarray(o.guest.sites, dtype=np.float64)])
    qs = np.asarray(o.guest.charges, dtype=np.float64)
    ph = ks @ sites.T
    sg = (np.cos(ph) * qs).sum(1) + 1j * (np.sin(ph) * qs).sum(1)
    return O.KE * float(np.sum(kp * np.abs(sg) ** 2))


def test_const_term_is_constant_offset_at_zero_and_exchange_constant_plus_self_mean():
    o = _system(tail=True)
    sc = O.lib_scalars(o)
    assert o.nvt_const_term(0) == sc[3]  # bit-identical, src/nvt.jl:48
    for n in (1, 2, 5, 17):
        ref = O.lib().or_exchange_constant(o._h, n) + sc[4]
        assert abs(o.nvt_const_term(n) - ref) < 1e-12 * (1 + abs(ref)), n
    # the Ng -> Ng+1 change is not the 0 -> 1 one when the tail or the net charge term is on
    assert o.nvt_const_term(4) != o.nvt_const_term(0)


@pytest.mark.parametrize("tail,charged", [(True, True), (False, True), (True, False)])
def test_chain_insertion_equals_total_energy_difference(tail, charged):
    o = _system(tail=tail, charged=charged)
    pos, quat = _poses(o, 9, seed=5)
    m = O.OracleMC(o, 12, 77, pos[:8], quat[:8])
    e0 = m.total_energy()
    mean_self = O.lib_scalars(o)[4]
    test_p, test_q = _poses(o, 6, seed=11, minsep=0.0, host_negative=False)
    for p, q in zip(test_p, test_q):
        dU = m.chain_insertion(p, q)
        m2 = O.OracleMC(o, 12, 77, np.vstack([pos[:8], p]), np.vstack([quat[:8], q]))
        exact = m2.total_energy() - e0
        corr = (_exact_self(o, q) - mean_self) if charged else 0.0
        assert abs(dU + corr - exact) <= 1e-9 * (1.0 + abs(exact)), (dU, corr, exact)


def test_chain_insertion_without_guests_is_the_widom_energy():
    o = _system(tail=True)
    m = O.OracleMC(o, 4, 5)
    seed = 2024
    dU_chain = m.chain_energies(seed, 1, 10, 40)
    dU_widom = o.energies(seed, 1, 10, 40)
    fin = np.isfinite(dU_widom) & (np.abs(dU_widom) < 1e6)
    assert fin.sum() > 20
    np.testing.assert_allclose(dU_chain[fin], dU_widom[fin], rtol=1e-11, atol=1e-11)


def test_nvt_without_guest_slots_reproduces_widom():
    """src/nvt.jl:182-210: N = 0 everywhere reproduces Widom on the same stream; energy 0, no
    move attempted."""
    o = _system(tail=True)
    chains = [O.OracleMC(o, 0, 11), O.OracleMC(o, 0, 12)]
    nw, NP = 6, 8
    res = O.nvt_run(chains, T=T, n_warmup=3, n_production=NP, n_widom_per_cycle=nw, n_audit=2, step_trans=0.3,
                    step_rot=0.3, widom_seed=42, nblocks=4)
    for i, r in enumerate(res):
        st, w, _ = o.widom(T, nw * NP, 42, i, 10)
        assert st == 0
        assert r["mu_ex"] == pytest.approx(w[0], rel=1e-10)
        assert r["K_H"] == pytest.approx(w[2], rel=1e-10)
        assert r["q_st"] == pytest.approx(w[4], rel=1e-10)
        assert r["energy"] == 0.0
        assert r["acceptance"] == (0.0, 0.0, 0.0) and int(np.sum(r["attempted"])) == 0


def test_nvt_cycle_length_is_shared_and_floored():
    """src/nvt.jl:212-253: one cycle length for the batch, max(max occupancy, min_cycle_length)."""
    o = _system(tail=False, natoms=20)
    pos, quat = _poses(o, 10, seed=9, minsep=3.2)
    chains = [O.OracleMC(o, 3, 21, pos[:3], quat[:3]), O.OracleMC(o, 7, 22, pos[3:10], quat[3:10])]
    res = O.nvt_run(chains, T=T, n_warmup=0, n_production=4, n_widom_per_cycle=2, n_audit=10**6, step_trans=0.3,
                    step_rot=0.3, widom_seed=1001, nblocks=2)
    assert [int(np.sum(r["attempted"])) for r in res] == [28, 28]
    chains = [O.OracleMC(o, 1, 31, pos[:1], quat[:1]), O.OracleMC(o, 1, 32, pos[1:2], quat[1:2])]
    res = O.nvt_run(chains, T=T, n_warmup=0, n_production=4, n_widom_per_cycle=1, n_audit=10**6, step_trans=0.3,
                    step_rot=0.3, min_cycle_length=5, widom_seed=1001, nblocks=2)
    assert [int(np.sum(r["attempted"])) for r in res] == [20, 20]


def test_nvt_samples_follow_the_chain():
    """Every production energy sample is the configuration's total energy, and every cycle's
    insertions are the chain insertions into the configuration of that cycle."""
    o = _system(tail=True)
    pos, quat = _poses(o, 5, seed=4)
    chain = O.OracleMC(o, 5, 99, pos, quat)
    twin = O.OracleMC(o, 5, 99, pos, quat)
    nw = 3
    res = O.nvt_run([chain], T=T, n_warmup=2, n_production=6, n_widom_per_cycle=nw, n_audit=4, step_trans=0.4,
                    step_rot=0.3, widom_seed=7, nblocks=3)[0]
    probs = (1.0, 1.0, 1.0, 0.0)
    since = 0
    for c in range(8):
        twin.cycle(probs, 5, 0.4, 0.3, 1.0, T)
        since += 1
        if since >= 4:
            twin.audit()
            since = 0
        if c >= 2:
            k = c - 2
            assert res["energies"][k] == pytest.approx(twin.total_energy(), rel=1e-9, abs=1e-9)
            np.testing.assert_allclose(res["dU"][k], twin.chain_energies(7, 0, k * nw, nw), rtol=1e-12, atol=1e-12)
    assert sum(res["accepted"]) > 0
