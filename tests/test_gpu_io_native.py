"""The C++ input readers feeding the device path: RUBTAK 2x2x2 + CO2 built from
pa_read_cif / pa_read_forcefield / pa_read_guest gives bit-identical Widom results to the same
files read by the Python readers (the parsed inputs are bit-equal, so every table is)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import pureadsorb_amd as pa
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_native_inputs_give_identical_widom_results():
    ew = pa.EwaldParams(12.0, 1e-6)
    ffn = pa.read_forcefield_native(os.path.join(GOLDEN, "trappe.yaml"))
    gn = pa.read_guest_native(os.path.join(GOLDEN, "co2.yaml"), ffn)
    fwn = pa.replicate(pa.read_cif_native(os.path.join(GOLDEN, "RUBTAK.cif")), (2, 2, 2))
    ffp = pa.read_forcefield(os.path.join(GOLDEN, "trappe.yaml"))
    gp = pa.read_guest(os.path.join(GOLDEN, "co2.yaml"), ffp)
    fwp = pa.replicate(pa.read_cif(os.path.join(GOLDEN, "RUBTAK.cif")), (2, 2, 2))
    for precision in ("f64", "mixed"):
        bn = pa.FrameworkBatch([fwn], ffn, gn, ew)
        bp = pa.FrameworkBatch([fwp], ffp, gp, ew)
        bn.bin_scale = bp.bin_scale = 1.0
        rn = pa.widom(bn, gn, T=298.15, ninsert=20000, seed=11, nblocks=4, precision=precision)[0]
        rp = pa.widom(bp, gp, T=298.15, ninsert=20000, seed=11, nblocks=4, precision=precision)[0]
        print(precision, rn.K_H, rp.K_H, flush=True)
        assert np.isfinite(rn.K_H) and rn.K_H > 0
        assert (rn.K_H, rn.q_st, rn.mu_ex) == (rp.K_H, rp.q_st, rp.mu_ex)
