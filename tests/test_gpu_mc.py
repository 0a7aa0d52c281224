"""GPU parity of the move-energy path (SystemState, moves, audit, GCMC driver) through the C ABI.

Bars: the device chain and the FP64 oracle chain (or_mc_*, the reference's move path restated
serially) start from the same poses and draw from the same counter-based streams, so every
attempt's slot and accept decision must agree exactly and its dU within 1e-9 relative (FP64 sums
in a different order); running energies, structure factors and poses likewise. Physics: the
grand-canonical ideal gas gives <N> = f V / kT (Loschmidt: 1 atm, 273.15 K, 37,219 A^3 -> 1)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from conftest import oracle_for
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu

T = 298.15
SEED = 5


def _porous_system(nsys=1):
    ff = synth.synthetic_forcefield(11.0)
    g = synth.co2_guest(ff)
    uc = synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5)
    fw = pa.replicate(uc, (2, 2, 2))
    ew = pa.EwaldParams(11.0)
    b = pa.FrameworkBatch([fw] * nsys, ff, g, ew, fullk=True)
    return fw, ff, g, ew, b


def _oracle_chain(state, n, fw, ff, g, ew, capacity):
    o = oracle_for(fw, ff, g, ew)
    p, q, _ = state.guests(n)
    return o, O.OracleMC(o, capacity, O.splitmix64(SEED, n + 1), p, q)


def test_state_creation_matches_oracle():
    fw, ff, g, ew, b = _porous_system()
    st = pa.SystemState(b, g, [8], T=T, seed=SEED, capacities=[40])
    o, m = _oracle_chain(st, 0, fw, ff, g, ew, 40)
    info = st.info(0)
    print(f"8 guests: device energy {info['energy']:.12f}, oracle {m.scalars()['energy']:.12f}", flush=True)
    assert info["occupancy"] == 8 and info["capacity"] == 40
    assert abs(info["energy"] - m.scalars()["energy"]) <= 1e-10 * (1 + abs(info["energy"]))
    _, _, he = st.guests(0)
    _, _, he_o = m.guests()
    assert np.allclose(he, he_o, rtol=1e-10, atol=1e-12)
    assert np.allclose(st.sk(0), m.sk(), rtol=0, atol=1e-11)
    assert abs(info["const_p"] - m.scalars()["const_p"]) < 1e-12
    assert abs(info["const_q"] - m.scalars()["const_q"]) < 1e-12
    assert np.allclose(pa.total_energy(st), m.total_energy(), rtol=1e-10)


@pytest.mark.parametrize("movetype,fug", [(1, 1e5), (2, 1e5), (3, 1e5), (4, 1e8), (5, 1e2)])
def test_move_traces_match_oracle(movetype, fug):
    fw, ff, g, ew, b = _porous_system()
    st = pa.SystemState(b, g, [8], T=T, seed=SEED, capacities=[60])
    o, m = _oracle_chain(st, 0, fw, ff, g, ew, 60)
    nm = 120
    tr = pa.mc_moves(st, movetype, nm, T=T, step_trans=0.3, step_rot=0.2, fugacity=fug, trace=True)[0]
    ref = np.array([m.attempt(movetype, 0.3, 0.2, fug, T) for _ in range(nm)])
    acc_dev = tr[:, 3].astype(int)
    print(f"type {movetype}: device accepted {acc_dev.sum()} / oracle {int(ref[:, 0].sum())} of {nm}", flush=True)
    assert np.array_equal(tr[:, 1].astype(int), ref[:, 2].astype(int))
    assert np.array_equal(acc_dev, ref[:, 0].astype(int))
    fin = np.isfinite(ref[:, 1])
    assert np.array_equal(np.isfinite(tr[:, 2]), fin)
    assert np.all(np.abs(tr[fin, 2] - ref[fin, 1]) <= 1e-9 * (1 + np.abs(ref[fin, 1])))
    info = st.info(0)
    assert info["occupancy"] == m.scalars()["occupancy"]
    assert abs(info["energy"] - m.scalars()["energy"]) <= 1e-9 * (1 + abs(info["energy"]))
    p, q, he = st.guests(0)
    po, qo, heo = m.guests()
    assert np.allclose(p, po, rtol=0, atol=1e-9) and np.allclose(q, qo, rtol=0, atol=1e-9)
    assert np.allclose(st.sk(0), m.sk(), rtol=0, atol=1e-9)
    d = pa.audit_energy(st)
    print(f"audit discrepancy {d[0]:.3e} eV", flush=True)
    assert abs(d[0]) < 1e-8


def test_mc_step_and_exchange_entry_points():
    fw, ff, g, ew, b = _porous_system(nsys=2)
    st = pa.SystemState(b, g, [4, 6], T=T, seed=SEED, capacities=[30, 30])
    for mt in (pa.MOVE_TRANSLATION, pa.MOVE_ROTATION, pa.MOVE_REINSERTION):
        for _ in range(10):
            pa.mc_step(st, mt, [0.5, 0.2], [0.3, 0.1], T=T)
    rng = np.random.default_rng(1)
    for _ in range(20):
        pa.mc_exchange(rng, st, [1e7, 1e6], T=T)
    for n in range(2):
        info = st.info(n)
        print(n, info, flush=True)
        assert sum(info["attempted"]) == 50 and info["rng_counter"] == 50
    e = pa.total_energy(st)
    assert np.allclose(e, st.energy(), rtol=1e-9, atol=1e-9)
    with pytest.raises(pa.ArgumentError, match="movetype"):
        pa.mc_step(st, 4, 0.1, 0.1, T=T)


def test_capacity_hit_is_an_error():
    fw, ff, g, ew, b = _porous_system()
    st = pa.SystemState(b, g, [0], T=T, seed=SEED, capacities=[1])
    with pytest.raises(pa.ArgumentError, match="capacity"):
        for _ in range(400):
            pa.mc_insert(st, [1e9], T=T)


def test_state_needs_fullk():
    fw, ff, g, ew, _ = _porous_system()
    b = pa.FrameworkBatch([fw], ff, g, ew)
    with pytest.raises(pa.ArgumentError, match="fullk"):
        pa.SystemState(b, g, [0], T=T)


def test_gcmc_driver_matches_oracle_cycles():
    fw, ff, g, ew, b = _porous_system()
    st = pa.SystemState(b, g, [3], T=T, seed=SEED, capacities=[80])
    o, m = _oracle_chain(st, 0, fw, ff, g, ew, 80)
    ncyc, f = 200, 2e6
    res = pa.run_gcmc(st, T=T, n_warmup=0, n_production=ncyc, n_audit=50, step_trans=0.5, step_rot=0.3,
                      fugacity=f, p_translation=0.25, p_rotation=0.25, p_exchange=0.5, min_cycle_length=2, nblocks=5)[0]
    e_dev, n_dev = pa.gcmc_samples(st, 0, ncyc)
    e_or, n_or = [], []
    for _ in range(ncyc):
        assert m.cycle((0.25, 0.25, 0.0, 0.5), 2, 0.5, 0.3, f, T) > 0
        s = m.scalars()
        e_or.append(s["energy"])
        n_or.append(s["occupancy"])
    print(f"loading {res.loading:.3f} +- {res.loading_err:.3f}, energy {res.energy:.4f}, q_st {res.q_st:.4f}, "
          f"oracle mean loading {np.mean(n_or):.3f}", flush=True)
    assert np.array_equal(n_dev, np.array(n_or, dtype=float))
    assert np.allclose(e_dev, e_or, rtol=1e-8, atol=1e-8)
    assert abs(res.loading - np.mean(n_or)) < 1e-12
    acc, att = m.counts()
    assert tuple(att) == res.attempted


def test_ideal_gas_loschmidt_device():
    L = 37219.0 ** (1 / 3)
    ff = pa.ForceField.from_types(["X_", "CH4"], [3.0, 3.73], [0.001, 148.0 * pa.KB], cutoff=12.0, tail=False)
    g = pa.Guest([[0.0, 0.0, 0.0]], [1], [0.0])
    fw = pa.Framework(np.diag([L, L, L]), np.zeros((0, 3)), [], [], [])
    b = pa.FrameworkBatch([fw], ff, g, pa.EwaldParams(12.0), fullk=True)
    st = pa.SystemState(b, g, [0], T=273.15, seed=3, capacities=[50])
    r = pa.run_gcmc(st, T=273.15, n_warmup=2000, n_production=40000, n_audit=10000, step_trans=1.0, step_rot=1.0,
                    fugacity=101325.0, p_translation=0.0, p_rotation=0.0, p_exchange=1.0, min_cycle_length=5,
                    nblocks=20)[0]
    print(f"<N> = {r.loading:.4f} +- {r.loading_err:.4f}", flush=True)
    assert abs(r.loading - 1.0) <= 4 * r.loading_err + 0.01
