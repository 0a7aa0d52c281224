"""GPU parity of phase 0 (the mixed path's hard-core rejection, src/widom.jl:99-151,202-253).

Bars: every insertion the device rejects is rejected by the oracle's brute-force check with
the reference's FP32 rounding allowance (the device's radii are never larger), and its FP64
oracle energy gives exp(-dU/kT) == 0.0 exactly; the two-phase and single-phase mixed paths
return identical Widom results (the reference's widom == widom_singlephase)."""
from __future__ import annotations

import numpy as np
import pytest

import pureadsorb_amd as pa
from conftest import oracle_for, oracle_threads
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu

T = 298.15
KT = pa.KB * T


def _porous(config1):
    _, ff, g, ew = config1
    uc = synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 80, 42, charge_sd=0.5)
    return pa.replicate(uc, (3, 3, 3)), ff, g, ew


@pytest.mark.parametrize("case", ["config1", "porous", "config0", "rubtak"])
def test_device_flags_against_oracle(case, config1, config0, rubtak, trappe, co2):
    if case == "config1":
        fw, ff, g, ew = config1
    elif case == "porous":
        fw, ff, g, ew = _porous(config1)
    elif case == "config0":
        fw, ff, g, ew = config0
    else:
        fw, ff, g, ew = pa.replicate(rubtak, (3, 3, 3)), trappe, co2, pa.EwaldParams(12.0)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    o = oracle_for(fw, ff, g, ew)
    n = 4000
    dev = pa.hardcore_flags(b, 0, 11, T, 0, n).astype(bool)
    _, rho2 = o.rejection_tables(KT, 1.1920929e-07)
    ref = o.hardcore_flags(11, 0, 0, n, rho2, nthreads=oracle_threads()).astype(bool)
    dU = o.energies(11, 0, 0, n, nthreads=oracle_threads())
    print(f"{case}: device rejects {dev.sum()}, oracle {ref.sum()} of {n}", flush=True)
    assert not np.any(dev & ~ref)
    assert dev.sum() >= 0.97 * ref.sum()
    assert np.all(np.exp(-dU[dev] / KT) == 0.0)


@pytest.mark.parametrize("case", ["porous", "rubtak", "config1"])
def test_two_phase_equals_single_phase(case, config1, rubtak, trappe, co2):
    if case == "porous":
        fw, ff, g, ew = _porous(config1)
    elif case == "config1":
        fw, ff, g, ew = config1
    else:
        fw, ff, g, ew = pa.replicate(rubtak, (3, 3, 3)), trappe, co2, pa.EwaldParams(12.0)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    r2 = pa.widom(b, g, T=T, ninsert=200_000, seed=5, nblocks=5)[0]
    r1 = pa.widom(b, g, T=T, ninsert=200_000, seed=5, nblocks=5, single_phase=True)[0]
    print(f"{case}: two-phase {r2}\nsingle-phase {r1}", flush=True)
    assert r2 == r1


def test_two_phase_batch_and_gradients(config1):
    fws = synth.config2_frameworks(256)
    sub = [fws[i] for i in (3, 40, 77, 200, 11, 150)]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    b = pa.FrameworkBatch(sub, ff, g, pa.EwaldParams(12.0))
    r2 = pa.widom(b, g, T=T, ninsert=6 * 20000, seed=8, nblocks=4)
    r1 = pa.widom(b, g, T=T, ninsert=6 * 20000, seed=8, nblocks=4, single_phase=True)
    for a, c in zip(r2, r1):
        print(a.ln_K_H, c.ln_K_H, flush=True)
        assert abs(a.ln_K_H - c.ln_K_H) <= 1e-12 * max(1.0, abs(c.ln_K_H))
        assert abs(a.q_st - c.q_st) <= 1e-12 * max(1.0, abs(c.q_st))
    g2 = pa.widom_grad(b, g, T=T, ninsert=6 * 20000, seed=8, nblocks=4)
    assert pa.last_timing(b)["grad_insertions"] > 0
    for s in range(len(sub)):
        assert g2[s].result.ln_K_H == r2[s].ln_K_H
