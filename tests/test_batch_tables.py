"""pa_batch_from_tables on the CPU: the argument checks it makes before touching a device, on
tables laid out as the reference's FrameworkBatch (src/batch.jl:72-104) from the oracle; a
valid set of tables reaches the device step (an error without a GPU, a batch with one)."""
from __future__ import annotations

import copy

import numpy as np
import pytest

import pureadsorb_amd as pa
from batch_tables import reference_tables


@pytest.fixture(scope="module")
def rubtak_tables(rubtak, trappe, co2):
    sc = pa.replicate(rubtak, (3, 3, 3))
    return reference_tables([sc, sc], trappe, co2, pa.EwaldParams(12.0, 1e-6)), co2


def test_layout_matches_the_reference_shapes(rubtak_tables):
    t, g = rubtak_tables
    assert list(t["framework_of"]) == [0, 0]
    assert t["atom_offsets"].tolist() == [0, 3078]
    assert t["k_offsets"].tolist() == [0, 190]  # coupled k-vectors of RUBTAK 3x3x3 (efficiency-design.md:13-14)
    assert t["kmin"].shape == (1, 3, len(t["compact_to_orig"]))
    assert np.all(t["kmin"] <= 0.0)


@pytest.mark.parametrize("mutate,exc,match", [
    (lambda t: t.__setitem__("atom_offsets", np.array([1, 3078], dtype=np.int32)), pa.DimensionMismatch, "start at 0"),
    (lambda t: t.__setitem__("framework_of", np.array([0, 1], dtype=np.int32)), pa.ArgumentError, "framework_of"),
    (lambda t: t["types"].__setitem__(7, 99), pa.ArgumentError, "compact LJ type"),
    (lambda t: t["ks"].__setitem__((3, 0), t["ks"][3, 0] + 0.01), pa.ArgumentError, "reciprocal-lattice vector"),
    (lambda t: t.__setitem__("cutoff", 20.0), pa.ArgumentError, "too small for cutoff"),
])
def test_invalid_tables_are_refused_before_any_device_work(rubtak_tables, mutate, exc, match):
    t, g = rubtak_tables
    t = copy.deepcopy(t)
    mutate(t)
    with pytest.raises(exc, match=match):
        pa.FrameworkBatch.from_tables(t, g)


def test_guest_type_outside_the_compact_types(rubtak_tables, trappe):
    t, g = rubtak_tables
    bad = pa.Guest(g.sites, [0, 0, 0], g.charges)  # a framework type the batch did not compact for the guest
    t = copy.deepcopy(t)
    t["compact_to_orig"] = np.array([x for x in t["compact_to_orig"] if x != 0], dtype=np.int32)
    n = len(t["compact_to_orig"])
    t["sigma"] = t["sigma"][:n, :n]
    t["epsilon"] = t["epsilon"][:n, :n]
    t["types"] = np.minimum(t["types"], n - 1)
    t["kmin"] = t["kmin"][:, :, :n]
    with pytest.raises(pa.ArgumentError, match="not among the batch's compact types"):
        pa.FrameworkBatch.from_tables(t, bad)


def test_valid_tables_reach_the_device_step(rubtak_tables):
    t, g = rubtak_tables
    try:
        b = pa.FrameworkBatch.from_tables(t, g)
    except pa.DeviceError as e:
        assert "no HIP device" in str(e)
        return
    assert b.nframeworks == 1 and b.nsys == 2 and b.natoms[0] == 3078 and b.nk[0] == 190
