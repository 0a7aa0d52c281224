"""Shared fixtures. GPU tests carry @pytest.mark.gpu; everything else runs on the CPU."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "pureadsorb.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import pureadsorb_amd as pa  # noqa: E402
from pureadsorb_amd import synth  # noqa: E402

import oracle as O  # noqa: E402  (the checker)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def oracle_threads() -> int:
    """Host threads for the oracle: at most the 16-core share of a GPU box."""
    return max(1, min(16, os.cpu_count() or 1))


def oracle_for(fw, ff, guest, ewald) -> "O.OracleSystem":
    """The oracle's system for one framework of a batch, with the same type mapping."""
    types = [pa.typeindex(ff, s + "_") for s in fw.symbols]
    return O.OracleSystem(fw.cell, fw.frac, types, fw.charges, fw.replication, ff.sigma, ff.epsilon, ff.cutoff,
                          ff.tail, guest.sites, guest.types, guest.charges, ewald.cutoff, ewald.precision)


@pytest.fixture(scope="session")
def rubtak():
    return pa.read_cif(os.path.join(GOLDEN, "RUBTAK.cif"))


@pytest.fixture(scope="session")
def trappe():
    return pa.read_forcefield(os.path.join(GOLDEN, "trappe.yaml"))


@pytest.fixture(scope="session")
def co2(trappe):
    return pa.read_guest(os.path.join(GOLDEN, "co2.yaml"), trappe)


@pytest.fixture(scope="session")
def config1():
    fw = synth.config1_framework()
    ff = synth.synthetic_forcefield(12.0)
    return fw, ff, synth.co2_guest(ff), pa.EwaldParams(12.0, 1e-6)


@pytest.fixture(scope="session")
def config0():
    fw = synth.config0_framework()
    ff = synth.synthetic_forcefield(12.8, guest="ch4")
    return fw, ff, synth.ch4_guest(ff), pa.EwaldParams(12.8, 1e-6)


def rel_close(a, b, rtol, floor=1.0):
    """|a - b| <= rtol * max(|b|, floor), elementwise, with matching infinities."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    same_inf = (np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b)))
    fin = np.isfinite(a) & np.isfinite(b)
    ok = same_inf | (fin & (np.abs(a - b) <= rtol * np.maximum(np.abs(b), floor)))
    return bool(np.all(ok)), ok
