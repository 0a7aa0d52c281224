"""Device-built FrameworkBatch tables (batch_kernels.hip: host structure factor and the
64-orientation guest self term on the device) against the host build of the same frameworks
(PA_BATCH_HOST_ONLY), which the CPU tests pin against the oracle. Bars: 1e-12 relative on the
hard-core bound, volume and alpha; 1e-14 eV absolute on constant_offset and the self-term
half-range, which are small differences of ~5e-3 eV terms (constant_offset of a neutral CO2 is
~1e-7 eV: the 64-orientation mean of the reciprocal self term cancels the Ewald self and
exclusion terms), so FP64 rounding of those terms (~1e-18 eV) is ~1e-10 of the result;
1e-10 * sum|q| on the structure factor (FP64 sums in another order)."""
from __future__ import annotations

import time

import numpy as np
import pytest

import pureadsorb_amd as pa
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu


def _compare(dev, host):
    for name in ("nk", "nk_full", "natoms"):
        assert np.array_equal(getattr(dev, name), getattr(host, name)), name
    for name in ("constant_offset", "self_term_halfrange"):
        err = float(np.max(np.abs(getattr(dev, name) - getattr(host, name))))
        print(f"{name}: max abs diff {err:.2e} eV", flush=True)
        assert err <= 1e-14, name
    for name in ("hardcore_bound", "volumes", "alphas"):
        d, h = getattr(dev, name), getattr(host, name)
        fin = np.isfinite(h)
        assert np.array_equal(np.isfinite(d), fin), name
        err = np.max(np.abs(d[fin] - h[fin]) / np.maximum(np.abs(h[fin]), 1e-300)) if fin.any() else 0.0
        print(f"{name}: max rel diff {err:.2e}", flush=True)
        assert err <= 1e-12, name


def test_device_tables_equal_host_tables_for_uneven_frameworks():
    fws = synth.config2_frameworks(256)[:24]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0, 1e-6)
    t0 = time.perf_counter()
    dev = pa.FrameworkBatch(fws, ff, g, ew)
    t1 = time.perf_counter()
    host = pa.FrameworkBatch(fws, ff, g, ew, host_only=True)
    t2 = time.perf_counter()
    print(f"24 frameworks: device build {t1 - t0:.2f} s, host build {t2 - t1:.2f} s", flush=True)
    _compare(dev, host)


def test_device_structure_factor_equals_host_through_the_move_path(rubtak, trappe, co2):
    """With fullk the state's structure factor with no guests is the host structure factor at
    every k-vector; the audit rebuilds it from the batch's host copy."""
    sc = pa.replicate(rubtak, (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    dev = pa.FrameworkBatch([sc], trappe, co2, ew, fullk=True)
    host = pa.FrameworkBatch([sc], trappe, co2, ew, fullk=True, host_only=True)
    _compare(dev, host)
    st = pa.SystemState(dev, co2, [0], T=298.15, seed=1)
    sk = st.sk(0)
    # the oracle's FP64 table for the same framework
    import oracle as O
    from conftest import oracle_for
    o = oracle_for(sc, trappe, co2, ew)
    _, _, S, _ = o.ktable(full=True)
    scale = float(np.sum(np.abs(sc.charges)))
    err = np.max(np.abs(sk - S)) / scale
    print(f"RUBTAK 3x3x3 full k table ({len(S)} k-vectors): max |dS| / sum|q| = {err:.2e}", flush=True)
    assert err <= 1e-10
    assert abs(pa.audit_energy(st)[0]) == 0.0
