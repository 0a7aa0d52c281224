"""pa_widom_noalloc (no device allocation; per-call buffers from one static device array) against
pa_widom on the same batch, seeds and insertion counts: every WidomResult field equal, on two
calls in a row, the second smaller. The bin scale is fixed first, because pa_widom may tune it
and pa_widom_noalloc never does. pa_widom launches the bins by descending count and
pa_widom_noalloc in spatial order, so equality also shows that the launch order changes no
energy."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import pureadsorb_amd as pa
from pureadsorb_amd import _lib
from pureadsorb_amd.batch import _GuestArgs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _fields(r):
    return (r.mu_ex, r.mu_ex_err, r.K_H, r.K_H_err, r.q_st, r.q_st_err, r.nsamples, r.nblocks, r.ln_K_H)


@pytest.fixture(scope="module")
def rubtak_batch():
    ff = pa.read_forcefield(os.path.join(GOLDEN, "trappe.yaml"))
    g = pa.read_guest(os.path.join(GOLDEN, "co2.yaml"), ff)
    sc = pa.replicate(pa.read_cif(os.path.join(GOLDEN, "RUBTAK.cif")), (3, 3, 3))
    b = pa.FrameworkBatch([sc, sc], ff, g, pa.EwaldParams(12.0, 1e-6))
    b.bin_scale = 1.0
    return b, g


@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_noalloc_equals_widom_on_two_calls(rubtak_batch, precision):
    b, g = rubtak_batch
    calls = [(400_000, 3), (60_000, 11)] if precision == "mixed" else [(40_000, 3), (6_000, 11)]
    for ninsert, seed in calls:
        print(f"no-alloc vs pa_widom: {precision}, {ninsert} insertions, seed {seed}", flush=True)
        ref = pa.widom(b, g, T=298.15, ninsert=ninsert, seed=seed, nblocks=5, precision=precision)
        got = pa.widom(b, g, T=298.15, ninsert=ninsert, seed=seed, nblocks=5, precision=precision, no_alloc=True)
        assert len(got) == len(ref) == 2
        for r, n in zip(ref, got):
            assert _fields(n) == _fields(r), (precision, ninsert, _fields(n), _fields(r))
            assert np.isfinite(n.ln_K_H)
        assert sum(n.nsamples for n in got) == ninsert


@pytest.mark.parametrize("kappa", [0.5, 2.0])
def test_both_launch_order_paths_match_the_spatial_order(rubtak_batch, kappa):
    """pa_widom orders the energy launch in one workgroup for grids of up to 14,336 bins and with
    hipCUB's radix sort above that. 200,000 insertions per system in two 36 A cells give
    2 x 26^3 = 35,152 bins at bin scale 0.5 (the sort) and 2 x 6^3 = 432 at 2.0 (one workgroup);
    pa_widom_noalloc takes the same bins in spatial order: every field equal."""
    b, g = rubtak_batch
    kept = b.bin_scale
    b.bin_scale = kappa
    try:
        print(f"launch order vs spatial order: bin scale {kappa}", flush=True)
        ref = pa.widom(b, g, T=298.15, ninsert=400_000, seed=7, nblocks=5, precision="mixed")
        got = pa.widom(b, g, T=298.15, ninsert=400_000, seed=7, nblocks=5, precision="mixed", no_alloc=True)
    finally:
        b.bin_scale = kept
    for r, n in zip(ref, got):
        assert _fields(n) == _fields(r)
        assert np.isfinite(n.ln_K_H)


def test_noalloc_single_phase_and_chunked(rubtak_batch):
    """Single-phase and a small chunk (several energy launches) through the fixed array."""
    b, g = rubtak_batch
    kw = dict(T=298.15, ninsert=100_000, seed=5, nblocks=4, precision="mixed", single_phase=True, chunk=30_000)
    ref = pa.widom(b, g, **kw)
    got = pa.widom(b, g, **kw, no_alloc=True)
    for r, n in zip(ref, got):
        assert _fields(n) == _fields(r)


def test_noalloc_rejects_a_call_larger_than_the_array(rubtak_batch):
    """1e7 insertions need 80 MB of energies alone: an argument error that states both sizes,
    raised before any launch; the next call still matches pa_widom."""
    b, g = rubtak_batch
    gs = _GuestArgs(g)
    out = (_lib.PaWidomResult * 2)()
    st = _lib.lib().pa_widom_noalloc(b.handle, C.byref(gs.s), 298.15, 10_000_000, 1, 5, 0, 0, _lib.PA_PREC_MIXED, out)
    assert st == _lib.PA_ERR_ARG
    msg = _lib.last_error()
    assert "no-allocation" in msg and str(64 << 20) in msg, msg
    ref = pa.widom(b, g, T=298.15, ninsert=20_000, seed=2, nblocks=4, precision="mixed")
    got = pa.widom(b, g, T=298.15, ninsert=20_000, seed=2, nblocks=4, precision="mixed", no_alloc=True)
    for r, n in zip(ref, got):
        assert _fields(n) == _fields(r)
