"""Pins the oracle's move-energy path (or_mc_*): the running energy accumulated from every
accepted move's dU equals a from-scratch total energy of the new configuration after every
attempt of every move type (the reference's own audit criterion, src/audit.jl), and a
grand-canonical ideal gas reproduces <N> = f V / kT (the Loschmidt check of
test/ideal_gas_tests.jl: 1 atm, 273.15 K, 37,219 A^3 -> <N> = 1)."""
import math

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from pureadsorb_amd import synth

from conftest import oracle_for

T = 298.15


def _system(tail=False, charged=True):
    ff = synth.synthetic_forcefield(11.0, tail=tail)
    g = synth.co2_guest(ff)
    if not charged:
        g = pa.Guest(g.sites, g.types, [0.0, 0.0, 0.0])
    uc = synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5)
    fw = pa.replicate(uc, (2, 2, 2))
    return fw, ff, g, oracle_for(fw, ff, g, pa.EwaldParams(11.0))


def _place(o, n, seed=3):
    """n guests with negative host energy and reference points >= 3.5 A apart."""
    rng = np.random.default_rng(seed)
    pos, quat = [], []
    while len(pos) < n:
        f = rng.random(3)
        p = o.cell @ f
        u = rng.random(3)
        q = np.array([math.sqrt(1 - u[0]) * math.cos(2 * math.pi * u[1]), math.sqrt(u[0]) * math.sin(2 * math.pi * u[2]),
                      math.sqrt(u[0]) * math.cos(2 * math.pi * u[2]), math.sqrt(1 - u[0]) * math.sin(2 * math.pi * u[1])])
        if O.lib().or_host_guest(o._h, np.ascontiguousarray(p), np.ascontiguousarray(q)) >= 0.0:
            continue
        if any(np.linalg.norm(p - x) < 3.5 for x in pos):
            continue
        pos.append(p)
        quat.append(q)
    return np.array(pos), np.array(quat)


@pytest.mark.parametrize("tail,charged", [(False, True), (True, True), (False, False)])
def test_running_energy_equals_total_after_every_attempt(tail, charged):
    fw, ff, g, o = _system(tail, charged)
    pos, quat = _place(o, 8)
    m = O.OracleMC(o, 40, 12345, pos, quat)
    assert abs(m.scalars()["energy"] - m.total_energy()) < 1e-12
    naccepted = 0
    for mt in (1, 2, 3, 4, 5, 1, 4, 2, 5, 3):
        for _ in range(40):
            st, dU, slot = m.attempt(mt, step_trans=0.3, step_rot=0.2, fugacity_pa=1e8 if mt == 4 else 1e2, T=T)
            assert st in (0, 1)
            naccepted += st
            e_run = m.scalars()["energy"]
            e_tot = m.total_energy()
            assert abs(e_run - e_tot) <= 1e-9 * (1.0 + abs(e_tot)), (mt, e_run, e_tot)
    acc, att = m.counts()
    print("accepted", acc, "attempted", att, "occupancy", m.scalars()["occupancy"])
    assert naccepted > 60 and np.all(acc[[0, 1, 3, 4]] > 0)


def test_exchange_constant_is_affine_in_N():
    fw, ff, g, o = _system(tail=True)
    c = [O.lib().or_exchange_constant(o._h, n) for n in range(6)]
    d = np.diff(c)
    assert np.allclose(d, d[0], rtol=1e-9, atol=1e-12)


def test_ideal_gas_loschmidt():
    """Empty box, one uncharged LJ site, exchange moves only: <N> = f V / kT = 1 at 1 atm,
    273.15 K, 37,219 A^3."""
    L = 37219.0 ** (1 / 3)
    ff = pa.ForceField.from_types(["X_", "CH4"], [3.0, 3.73], [0.001, 148.0 * pa.KB], cutoff=12.0, tail=False)
    g = pa.Guest([[0.0, 0.0, 0.0]], [1], [0.0])
    fw = pa.Framework(np.diag([L, L, L]), np.zeros((0, 3)), [], [], [])
    o = oracle_for(fw, ff, g, pa.EwaldParams(12.0))
    Tg = 273.15
    f = 101325.0
    expect = f * O.PASCAL * 37219.0 / (O.KB * Tg)
    m = O.OracleMC(o, 50, 777)
    for _ in range(2000):
        m.cycle((0, 0, 0, 1), 5, 0.0, 0.0, f, Tg)
    n = []
    for _ in range(40000):
        m.cycle((0, 0, 0, 1), 5, 0.0, 0.0, f, Tg)
        n.append(m.scalars()["occupancy"])
    n = np.array(n, dtype=float)
    blocks = n.reshape(20, -1).mean(axis=1)
    mean, sem = blocks.mean(), blocks.std(ddof=1) / math.sqrt(len(blocks))
    print(f"<N> = {mean:.4f} +- {sem:.4f}, expected {expect:.4f}")
    assert abs(expect - 1.0) < 1e-3
    assert abs(mean - expect) <= 4 * sem + 0.01
