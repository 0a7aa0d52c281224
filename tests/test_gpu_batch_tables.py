"""pa_batch_from_tables on the device: a batch taken over from tables laid out as the reference's
FrameworkBatch (src/batch.jl:72-104, every table from the FP64 oracle, atoms cell-sorted) gives
the same Widom results as the batch pa_batch_create builds from the Frameworks, to rounding
(summation order and the 1-ulp Cartesian/fractional round trip), on both precisions; the
reference's B_s (exact per-atom terms) only changes which zero-weight insertions phase 0
rejects."""
from __future__ import annotations

import numpy as np
import pytest

import pureadsorb_amd as pa
from batch_tables import reference_tables
from conftest import rel_close
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu

T = 298.15


def _same_results(b_tab, b_own, g, ninsert, seed, prec, tol):
    rt = pa.widom(b_tab, g, T=T, ninsert=ninsert, seed=seed, nblocks=5, precision=prec)
    ro = pa.widom(b_own, g, T=T, ninsert=ninsert, seed=seed, nblocks=5, precision=prec)
    for a, b in zip(rt, ro):
        print(f"{prec}: ln K_H tables {a.ln_K_H:.12f} own {b.ln_K_H:.12f}; q_st {a.q_st:.12f} {b.q_st:.12f}", flush=True)
        assert abs(a.ln_K_H - b.ln_K_H) <= tol * max(1.0, abs(b.ln_K_H))
        assert abs(a.q_st - b.q_st) <= tol * max(1.0, abs(b.q_st))
        assert abs(a.mu_ex - b.mu_ex) <= tol * max(1.0, abs(b.mu_ex))
        assert a.nsamples == b.nsamples


def test_rubtak_from_tables_matches_pa_batch_create(rubtak, trappe, co2):
    sc = pa.replicate(rubtak, (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    print("RUBTAK 3x3x3 + CO2 from FrameworkBatch tables", flush=True)
    bt = pa.FrameworkBatch.from_tables(reference_tables([sc], trappe, co2, ew), co2)
    bo = pa.FrameworkBatch([sc], trappe, co2, ew)
    assert bt.nk[0] == bo.nk[0] == 190 and bt.natoms[0] == 3078
    assert bt.constant_offset[0] == pytest.approx(bo.constant_offset[0], abs=1e-12)
    bt.bin_scale = bo.bin_scale = 1.0
    _same_results(bt, bo, co2, 200_000, 42, "mixed", 1e-8)
    _same_results(bt, bo, co2, 20_000, 42, "f64", 1e-10)
    e_t = pa.widom_energies(bt, 0, 7, 0, 4096, precision="f64")
    e_o = pa.widom_energies(bo, 0, 7, 0, 4096, precision="f64")
    ok, _ = rel_close(e_t, e_o, 1e-11)
    assert ok
    # gradients and chains need the force field behind the tables
    with pytest.raises(pa.ArgumentError, match="pa_batch_create"):
        pa.widom_grad(bt, co2, T=T, ninsert=2000, seed=1, nblocks=4)
    bk = pa.FrameworkBatch.from_tables(reference_tables([sc], trappe, co2, ew, fullk=True), co2)
    assert bk.nk[0] == 4587
    with pytest.raises(pa.ArgumentError, match="pa_batch_create"):
        pa.SystemState(bk, co2, [1], T=T)


def test_multi_framework_tables_match():
    """Three systems on two distinct config-2 frameworks (system-major offsets, framework_of)
    and the uncharged CH4 case (no k table)."""
    fws = synth.config2_frameworks(256)
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    sel = [fws[40], fws[42], fws[40]]
    bt = pa.FrameworkBatch.from_tables(reference_tables(sel, ff, g, ew), g)
    bo = pa.FrameworkBatch(sel, ff, g, ew)
    assert list(bt.framework_of) == [0, 1, 0] and list(bo.framework_of) == [0, 1, 0]
    bt.bin_scale = bo.bin_scale = 1.0
    _same_results(bt, bo, g, 300_000, 5, "mixed", 1e-8)
    _same_results(bt, bo, g, 30_000, 5, "f64", 1e-10)

    fw0 = synth.config0_framework()
    ff0 = synth.synthetic_forcefield(12.8, guest="ch4")
    g0 = synth.ch4_guest(ff0)
    ew0 = pa.EwaldParams(12.8, 1e-6)
    bt0 = pa.FrameworkBatch.from_tables(reference_tables([fw0], ff0, g0, ew0), g0)
    bo0 = pa.FrameworkBatch([fw0], ff0, g0, ew0)
    assert bt0.nk[0] == 0
    bt0.bin_scale = bo0.bin_scale = 1.0
    _same_results(bt0, bo0, g0, 100_000, 3, "mixed", 1e-8)
    _same_results(bt0, bo0, g0, 20_000, 3, "f64", 1e-10)
