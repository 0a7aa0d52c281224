"""Pins the oracle's hard-core rejection (src/reject.jl; src/widom.jl:99-151,202-253): theta_F is
the exact underflow point of exp, the bracketed roots are sign changes of their functions, and
the property the construction exists for -- every insertion phase 0 flags has a Boltzmann
weight of exactly 0.0 in Float64 -- holds for the oracle's own FP64 energies on random poses."""
import math

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from pureadsorb_amd import synth

from conftest import oracle_for, oracle_threads

T = 298.15
KT = pa.KB * T


def test_theta_is_the_exp_underflow_point():
    th = O.theta_f64()
    assert math.exp(-th) == 0.0
    assert math.exp(-math.nextafter(th, 0.0)) > 0.0
    assert 745.0 < th < 745.2


@pytest.mark.parametrize("K", [-0.5, -3.0, -20.0])
def test_find_r0_is_the_inner_zero(K):
    sg, ep, alpha = 3.1, 0.004, 0.2667
    r0 = O.find_r0(sg, ep, K, alpha, 12.0)
    assert r0 < sg * 2 ** (1 / 6)
    assert O.pair_energy(r0 * (1 - 1e-9), sg, ep, K, alpha) > 0.0
    assert O.pair_energy(r0 * (1 + 1e-9), sg, ep, K, alpha) < 0.0
    # no other zero below r0: the pair energy only grows toward r -> 0
    rs = np.linspace(0.05, r0 * (1 - 1e-6), 400)
    assert all(O.pair_energy(r, sg, ep, K, alpha) > 0.0 for r in rs)


def test_find_rho2_solves_its_equation_and_clamps():
    sg, ep, kmin, margin = 3.0, 0.005, -4.0, 50.0
    rho = math.sqrt(O.find_rho2(sg, ep, kmin, margin, 12.0))

    def f(r):
        x = (sg / r) ** 6
        return 4 * ep * (x * x - x) - abs(kmin) / r - margin
    assert f(rho * (1 - 1e-9)) > 0.0 > f(rho * (1 + 1e-9))
    assert O.find_rho2(sg, ep, kmin, math.inf, 12.0) == 0.0
    assert O.find_rho2(sg, 0.0, kmin, margin, 12.0) == 0.0
    # a huge well (tiny margin relative to the LJ wall) clamps at the LJ cutoff
    assert O.find_rho2(sg, 1e9, 0.0, 1e-12, 2.5) == 2.5 ** 2


@pytest.mark.parametrize("case", ["config1", "porous", "uncharged"])
def test_flagged_insertions_have_exactly_zero_weight(case, config1, config0):
    if case == "config1":
        fw, ff, g, ew = config1
    elif case == "porous":
        _, ff, g, ew = config1
        uc = synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 80, 42, charge_sd=0.5)
        fw = pa.replicate(uc, (3, 3, 3))
    else:
        fw, ff, g, ew = config0
    o = oracle_for(fw, ff, g, ew)
    for eps_f in (2.220446049250313e-16, 1.1920929e-07):
        margin, rho2 = o.rejection_tables(KT, eps_f)
        assert margin > 0 and np.all(rho2 >= 0) and np.all(rho2 < ff.cutoff ** 2)
        n = 3000
        flags = o.hardcore_flags(7, 0, 0, n, rho2, nthreads=oracle_threads())
        dU = o.energies(7, 0, 0, n, nthreads=oracle_threads())
        print(f"{case} eps_f={eps_f:.3g}: B_s {o.hardcore_bound():.6g} eV, margin {margin:.6g} eV, "
              f"rho max {math.sqrt(rho2.max()):.4f} A, flagged {flags.sum()} of {n}, "
              f"min flagged dU/kT {dU[flags == 1].min() / KT if flags.any() else float('nan'):.1f}")
        assert np.all(np.exp(-dU[flags == 1] / KT) == 0.0)
        if case == "config1":
            assert flags.sum() > n // 4  # the dense framework rejects a large share
