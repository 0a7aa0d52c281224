"""Host side of the product on the CPU (no device): the C ABI library loads and exports what
include/pureadsorb_amd.h declares; FrameworkBatch tables built host-only agree with the
oracle; argument checks keep the reference's wording; CIF/YAML input and synthetic inputs."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from conftest import GOLDEN, ROOT, oracle_for
from pureadsorb_amd import _lib, synth


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "pureadsorb_amd.h")) as fh:
        declared = set(re.findall(r"\b(pa_[a-z_0-9]+)\s*\(", fh.read()))
    assert declared == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_product_coefficients_equal_the_pinned_literals():
    """The device series (pa_numeric.hpp) uses exactly the literals the oracle's regeneration
    test pins."""
    src = open(os.path.join(ROOT, "pureadsorb.jl_amd", "csrc", "pa_numeric.hpp")).read()
    block = src[src.index("inline double pair_erfc_f64_t"):]
    block = block[block.index("{", block.index("c[17]")) + 1: block.index("};")]
    vals = np.array([float(x) for x in re.findall(r"[-+]?\d[\d.]*e?[-+]?\d*", block.replace("\n", " "))])
    assert np.array_equal(vals, O.erfc_coefficients("pair64"))
    f32 = src[src.index("pair_erfc_c32"):]
    vals32 = np.array([float(x.rstrip("f")) for x in re.findall(r"[-+]?\d[\d.]*(?:e[-+]?\d+)?f", f32)[:8]], dtype=np.float32)
    assert np.array_equal(vals32, O.erfc_coefficients("pair32").astype(np.float32))


def _host_batch(fw, ff, g, ew, **kw):
    return pa.FrameworkBatch([fw], ff, g, ew, host_only=True, **kw)


@pytest.mark.parametrize("case", ["config0", "config1", "rubtak", "rubtak_fullk"])
def test_host_tables_match_oracle(case, config0, config1, rubtak, trappe, co2):
    if case == "config0":
        fw, ff, g, ew = config0
    elif case == "config1":
        fw, ff, g, ew = config1
    else:
        fw, ff, g, ew = pa.replicate(rubtak, (3, 3, 3)), trappe, co2, pa.EwaldParams(12.0, 1e-6)
    b = _host_batch(fw, ff, g, ew, fullk=case.endswith("fullk"))
    o = oracle_for(fw, ff, g, ew)
    assert b.natoms[0] == fw.natoms
    assert b.nk[0] == (o.nkfull if case.endswith("fullk") else o.nk)
    assert b.nk_full[0] == o.nkfull
    assert b.volumes[0] == pytest.approx(o.V, rel=1e-14)
    assert b.alphas[0] == o.alpha
    assert b.constant_offset[0] == pytest.approx(o.constant_offset, rel=1e-11, abs=1e-14)
    assert b.self_term_halfrange[0] == pytest.approx(o.self_halfrange, rel=1e-9, abs=1e-18)


def test_tail_correction_in_constant_offset(rubtak, trappe, co2):
    """trappe.yaml turns the tail correction on: constant_offset carries tail_delta
    (src/forcefield.jl:111-124)."""
    sc = pa.replicate(rubtak, (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    b = _host_batch(sc, trappe, co2, ew)
    o = oracle_for(sc, trappe, co2, ew)
    counts = np.bincount([trappe.names.index(s + "_") for s in sc.symbols], minlength=trappe.ntypes)
    gcounts = np.bincount(co2.types, minlength=trappe.ntypes)
    tail = pa.tail_delta(trappe, counts, gcounts, pa.volume(sc.cell))
    assert tail != 0.0 and o.tail_term == pytest.approx(tail, rel=1e-12)
    assert b.constant_offset[0] == pytest.approx(o.constant_offset, rel=1e-11)


def test_dedup_by_value(config1):
    fw, ff, g, ew = config1
    copy = pa.Framework(fw.cell.copy(), fw.frac.copy(), fw.labels, fw.symbols, fw.charges.copy(), fw.replication)
    b = pa.FrameworkBatch([fw, copy, fw], ff, g, ew, host_only=True)
    assert b.nsys == 3 and b.nframeworks == 1 and list(b.framework_of) == [0, 0, 0]


def test_neutral_guest_builds_no_k_table(config0):
    fw, ff, g, ew = config0
    b = _host_batch(fw, ff, g, ew)
    assert b.nk[0] == 0 and b.self_term_halfrange[0] == 0.0


def test_batch_errors_keep_reference_wording(rubtak, trappe, co2, config1):
    with pytest.raises(pa.ArgumentError, match="replicate"):
        _host_batch(rubtak, trappe, co2, pa.EwaldParams(12.0))
    sc = pa.replicate(rubtak, (3, 3, 3))
    with pytest.raises(pa.ArgumentError, match="exceeds PAIR_ERFC_XMAX"):
        _host_batch(sc, trappe, co2, pa.EwaldParams(12.0, 1e-12))
    with pytest.raises(pa.ArgumentError, match="loosen ewald.precision or shorten ewald.cutoff"):
        _host_batch(sc, trappe, co2, pa.EwaldParams(12.0, 1e-12))
    fw, ff, g, ew = config1
    lie = pa.Framework(fw.cell, fw.frac, fw.labels, fw.symbols, np.random.default_rng(1).normal(size=fw.natoms), (3, 3, 3))
    with pytest.raises(pa.ArgumentError, match="does not describe the atoms"):
        _host_batch(lie, ff, g, ew)
    bad = pa.Framework(fw.cell, fw.frac, fw.labels, ["Xx"] * fw.natoms, fw.charges, fw.replication)
    with pytest.raises(pa.ArgumentError, match="unknown LJ type"):
        _host_batch(bad, ff, g, ew)


def test_list_capacity_knob_keeps_every_pass_progressing(config1):
    """The mixed path's neighbour list is built in slices of 768 atoms (one per work-item of a bin
    workgroup); a capacity below one slice could never take a slice and would never finish, so
    the knob refuses it, as it refuses a capacity that is not whole 64-entry rows or exceeds the
    LDS budget. Before any mixed call the statistics report nothing."""
    fw, ff, g, ew = config1
    b = _host_batch(fw, ff, g, ew)
    for bad in (64, 448, 700, 800, 1 << 20, -64):
        with pytest.raises(pa.ArgumentError, match="list capacity"):
            b.list_capacity = bad
    b.list_capacity = 768
    assert b.list_capacity == 768
    b.list_capacity = 0
    with pytest.raises(pa.ArgumentError, match="HOST_ONLY"):
        b.list_stats()


def test_widom_argument_checks_keep_reference_wording(config1):
    """Checked before any device work (src/widom.jl:302-330; test/widom_tests.jl:183-230)."""
    fw, ff, g, ew = config1
    b = _host_batch(fw, ff, g, ew)
    with pytest.raises(pa.ArgumentError, match="ninsert"):
        pa.widom(b, g, T=300.0, ninsert=5, nblocks=10)
    with pytest.raises(pa.ArgumentError, match="nblocks"):
        pa.widom(b, g, T=300.0, ninsert=100, nblocks=1)
    with pytest.raises(pa.ArgumentError, match="chunk"):
        pa.widom(b, g, T=300.0, ninsert=100, chunk=0)
    with pytest.raises(pa.ArgumentError, match="run=0 must be ≥ 1"):
        pa.widom(b, g, T=300.0, ninsert=100, run=0)
    with pytest.raises(pa.ArgumentError, match="run=1000 must be ≤ per=100"):
        pa.widom(b, g, T=300.0, ninsert=100, run=1000)
    other = pa.Guest(g.sites, g.types, g.charges * 2)
    with pytest.raises(pa.ArgumentError, match="does not match the guest"):
        pa.widom(b, other, T=300.0, ninsert=100)
    # arguments that pass every check reach the device stage, which a host-only batch refuses
    with pytest.raises(pa.ArgumentError, match="HOST_ONLY"):
        pa.widom(b, g, T=300.0, ninsert=100)


def test_read_cif_rubtak(rubtak):
    """test/structure_tests.jl:1-12"""
    assert rubtak.natoms == 114 and rubtak.symbols[0] == "Zr"
    assert rubtak.charges[0] == pytest.approx(2.38565)
    assert np.allclose(rubtak.frac[0], [0.37986, 0.37998, 0.61969])
    assert abs(rubtak.total_charge()) < 1e-3
    assert np.allclose(rubtak.cell, pa.cell_matrix(14.7619, 14.80147, 14.76539, 59.84578, 60.04729, 59.8131))
    assert rubtak.replication == (1, 1, 1)


def test_replicate(rubtak):
    """test/structure_tests.jl:14-28"""
    sc = pa.replicate(rubtak, (3, 3, 3))
    assert sc.natoms == 27 * 114
    assert np.allclose(sc.cell, 3 * rubtak.cell)
    assert sc.total_charge() == pytest.approx(27 * rubtak.total_charge(), abs=1e-9)
    assert np.all((sc.frac >= 0) & (sc.frac < 1))
    assert sc.replication == (3, 3, 3)
    assert pa.replicate(pa.replicate(rubtak, (2, 1, 1)), (1, 3, 2)).replication == (2, 3, 2)
    # image (i, j, k) with i fastest, atoms innermost (src/structure.jl:79-92)
    assert np.allclose(sc.frac[114], (rubtak.frac[0] + [1, 0, 0]) / 3)


def test_read_cif_rejects_non_p1_and_chargeless(tmp_path):
    p = tmp_path / "bad.cif"
    p.write_text("data_x\n_symmetry_space_group_name_H-M 'P 21'\nloop_\n_atom_site_label\n_atom_site_fract_x\n")
    with pytest.raises(pa.ArgumentError, match="P1"):
        pa.read_cif(str(p))
    src = open(os.path.join(GOLDEN, "RUBTAK.cif")).read()
    p.write_text(src.replace("_atom_site_charge\n", ""))
    with pytest.raises(pa.ArgumentError, match="_atom_site_charge"):
        pa.read_cif(str(p))


def test_forcefield_and_guest_input(trappe, co2):
    """test/forcefield_tests.jl:1-33"""
    ff = pa.ForceField.from_types(["A", "B"], [3.0, 4.0], [0.01, 0.04], cutoff=12.0)
    assert ff.sigma[0, 1] == pytest.approx(3.5) and ff.epsilon[0, 1] == pytest.approx(0.02)
    assert pa.typeindex(ff, "B") == 1
    with pytest.raises(pa.ArgumentError, match="unknown LJ type"):
        pa.typeindex(ff, "C")
    assert pa.tail_coefficient(ff, 0, 1) == pytest.approx(0.02 * 3.5**3 * ((3.5 / 12) ** 9 / 3 - (3.5 / 12) ** 3))
    c = lambda i, j: pa.tail_coefficient(ff, i, j)  # noqa: E731
    assert pa.tail_delta(ff, [10, 0], [0, 2], 1000.0) == pytest.approx(8 * np.pi / 3 / 1000 * (2 * (10 * 2 * c(0, 1)) + 4 * c(1, 1)))
    assert trappe.cutoff == 12.0 and trappe.tail
    i = pa.typeindex(trappe, "Ac_")
    assert trappe.sigma[i, i] == pytest.approx(3.0985) and trappe.epsilon[i, i] == pytest.approx(0.0014311662224050347)
    assert co2.nsites == 3 and list(co2.charges) == [0.7, -0.35, -0.35]
    assert np.allclose(co2.sites[1], [-1.16, 0, 0]) and co2.tc == pytest.approx(303.75) and co2.pc == 7.84e6


def test_cell_matrix_matches_oracle():
    for p in ((13, 13, 15, 90, 90, 120), (14.7619, 14.80147, 14.76539, 59.84578, 60.04729, 59.8131), (10, 20, 30, 90, 90, 90)):
        A = pa.cell_matrix(*p)
        assert np.allclose(A, O.cell_matrix(*p), rtol=1e-15, atol=1e-14)
    A = pa.cell_matrix(12.9, 12.9, 12.9, 90, 90, 90)
    assert np.count_nonzero(A - np.diag(np.diag(A))) == 0


def test_sys_of_index_and_counts():
    for ninsert, nsys, run in ((37, 4, 3), (100, 7, 5), (11, 3, 2), (1, 1, 1), (500, 13, 1), (45, 2, 5)):
        counts = np.zeros(nsys, dtype=np.int64)
        for g in range(1, ninsert + 1):
            s = pa.sys_of_index(g, run, nsys)
            assert 1 <= s <= nsys
            counts[s - 1] += 1
        assert np.array_equal(counts, pa.system_counts(ninsert, nsys, run))
        assert np.array_equal(pa.system_counts(ninsert, nsys, run), O.system_counts(ninsert, nsys, run))


def test_synthetic_frameworks_follow_their_rule():
    fw = synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 300, 42, charge_sd=0.5)
    again = synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 300, 42, charge_sd=0.5)
    assert np.array_equal(fw.frac, again.frac) and np.array_equal(fw.charges, again.charges)
    assert fw.natoms == 300 and abs(fw.total_charge()) < 1e-12
    assert set(fw.symbols) <= {"C", "H", "O", "Zn"}
    d = fw.frac[:, None, :] - fw.frac[None, :, :]
    d -= np.rint(d)
    r = np.linalg.norm(d @ fw.cell.T, axis=2) + np.eye(fw.natoms) * 99
    assert r.min() >= 1.5
    assert 0.3 < fw.charges.std() < 0.7
    # a count above 0.15 atoms per cubic angstrom is capped
    dense = synth.synthetic_framework(10.0, 10.0, 10.0, 90, 90, 90, 6000, 1)
    assert dense.natoms == 150
    c1 = synth.config1_framework()
    assert c1.natoms == 8100 and c1.replication == (3, 3, 3)


def test_grad_rejects_non_lb_tables_and_shuffled_images():
    """pa_widom_grad's own checks run before any device work (host-only batch)."""
    from pureadsorb_amd import synth
    ff = synth.synthetic_forcefield(11.0)
    g = synth.co2_guest(ff)
    uc = synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5)
    fw = pa.replicate(uc, (2, 2, 2))
    b = pa.FrameworkBatch([fw], ff, g, pa.EwaldParams(11.0), host_only=True)
    with pytest.raises(pa.ArgumentError, match="HOST_ONLY"):
        pa.widom_grad(b, g, T=298.15, ninsert=1000, nblocks=4)
    # atoms of the replicated cell out of replicate order: the claim still holds, the image map does not
    perm = np.random.default_rng(0).permutation(fw.natoms)
    shuffled = pa.Framework(fw.cell, fw.frac[perm], [fw.labels[i] for i in perm], [fw.symbols[i] for i in perm],
                            np.asarray(fw.charges)[perm], fw.replication)
    b2 = pa.FrameworkBatch([shuffled], ff, g, pa.EwaldParams(11.0), host_only=True)
    with pytest.raises(pa.ArgumentError, match="replicate order"):
        pa.widom_grad(b2, g, T=298.15, ninsert=1000, nblocks=4)
    # a force field whose cross terms are not Lorentz-Berthelot
    sig = np.array(ff.sigma, dtype=np.float64)
    sig[0, 1] = sig[1, 0] = sig[0, 1] * 1.1
    ff2 = pa.ForceField(ff.names, sig, np.array(ff.epsilon), ff.cutoff, ff.tail)
    b3 = pa.FrameworkBatch([fw], ff2, g, pa.EwaldParams(11.0), host_only=True)
    with pytest.raises(pa.ArgumentError, match="Lorentz-Berthelot"):
        pa.widom_grad(b3, g, T=298.15, ninsert=1000, nblocks=4)


def test_hardcore_bound_never_below_the_reference_bound(config1):
    """B_s with the per-atom terms rounded up on the |K| grid: never below the oracle's exact
    reference bound, and within the grid's 2 % of it."""
    import oracle as O
    from conftest import oracle_for
    fw, ff, g, ew = config1
    b = pa.FrameworkBatch([fw], ff, g, ew, host_only=True)
    ref = oracle_for(fw, ff, g, ew).hardcore_bound()
    print(b.hardcore_bound[0], ref)
    assert ref <= b.hardcore_bound[0] <= 1.021 * ref
