"""GPU parity of the gradient path (pa_widom_grad through the C ABI).

Bars: against the FP64 oracle gradient (or_widom_grad, all insertions, FP64 pair terms) and
against central finite differences of the device's own FP64 Widom on common random numbers
(relative step 1e-4), every per-type entry within 1e-3 relative; charge entries within 1e-3 of
the largest entry of their system.
"""
from __future__ import annotations

import numpy as np
import pytest

import pureadsorb_amd as pa
from conftest import oracle_for, oracle_threads
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu

T = 298.15
TYPES = [("C_", 52.8, 3.43), ("H_", 22.1, 2.57), ("O_", 30.2, 3.12), ("Zn_", 62.4, 2.46), ("C_co2", 27.0, 2.80),
         ("O_co2", 79.0, 3.05)]
RTOL = 1e-3


def _ff(eps_scale=None, sig_scale=None, tail=False):
    names = [t[0] for t in TYPES]
    eps = np.array([t[1] * pa.KB for t in TYPES])
    sig = np.array([t[2] for t in TYPES])
    if eps_scale is not None:
        eps = eps * eps_scale
    if sig_scale is not None:
        sig = sig * sig_scale
    return pa.ForceField.from_types(names, list(sig), list(eps), cutoff=11.0, tail=tail)


@pytest.fixture(scope="module")
def unit_cell():
    return synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5)


def _close_types(an, ref, label):
    for t in range(len(ref)):
        print(f"{label}[{t}]: device {an[t]:.8g} oracle {ref[t]:.8g}", flush=True)
    assert np.all(np.abs(an - ref) <= RTOL * np.abs(ref) + 1e-9), label


def _close_atoms(an, ref, label):
    err = np.abs(an - ref).max()
    print(f"{label}: max |d| {err:.3g} of max |ref| {np.abs(ref).max():.3g}", flush=True)
    assert err <= RTOL * np.abs(ref).max(), label


@pytest.mark.parametrize("precision,tail", [("f64", False), ("mixed", False), ("mixed", True)])
def test_grad_matches_oracle_gradient(unit_cell, precision, tail):
    fw = pa.replicate(unit_cell, (2, 2, 2))
    ff = _ff(tail=tail)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(11.0)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    ns = 20000
    print(f"grad vs oracle: {fw.natoms} atoms, {ns} insertions, {precision}, tail={tail}", flush=True)
    gd = pa.widom_grad(b, g, T=T, ninsert=ns, seed=3, nblocks=4, precision=precision)[0]
    ref = oracle_for(fw, ff, g, ew).widom_grad(T, ns, 3, 0, nthreads=oracle_threads())
    assert abs(gd.result.ln_K_H - ref["ln_K_H"]) <= 1e-4
    for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
        _close_types(getattr(gd, key), ref[key], key)
    for key in ("dlnkh_dq", "dqst_dq"):
        _close_atoms(getattr(gd, key), ref[key], key)
    # every replication image reports the same entry
    v = gd.dlnkh_dq.reshape(8, -1)
    assert np.allclose(v, v[0], rtol=1e-12, atol=0)
    t = pa.last_timing(b)
    print(f"grad kernel {t['grad_ms']:.3f} ms for {t['grad_insertions']} weighted insertions", flush=True)


def _lnk_qst_dev(fw, ff, ns, seed):
    g = synth.co2_guest(ff)
    b = pa.FrameworkBatch([fw], ff, g, pa.EwaldParams(11.0))
    r = pa.widom(b, g, T=T, ninsert=ns, seed=seed, nblocks=4, precision="f64")[0]
    return r.ln_K_H, r.q_st


def test_grad_matches_device_finite_differences(unit_cell):
    """Central differences of the device FP64 Widom (relative step 1e-4, same insertions)."""
    fw = pa.replicate(unit_cell, (2, 2, 2))
    ff = _ff()
    g = synth.co2_guest(ff)
    b = pa.FrameworkBatch([fw], ff, g, pa.EwaldParams(11.0))
    ns, seed, h = 50000, 9, 1e-4
    gd = pa.widom_grad(b, g, T=T, ninsert=ns, seed=seed, nblocks=4, precision="f64")[0]
    for t in range(4):
        for kind in ("eps", "sig"):
            up, dn = np.ones(len(TYPES)), np.ones(len(TYPES))
            up[t] += h
            dn[t] -= h
            key = "eps_scale" if kind == "eps" else "sig_scale"
            lp, qp = _lnk_qst_dev(fw, _ff(**{key: up}), ns, seed)
            lm, qm = _lnk_qst_dev(fw, _ff(**{key: dn}), ns, seed)
            base = TYPES[t][1] * pa.KB if kind == "eps" else TYPES[t][2]
            fd_l, fd_q = (lp - lm) / (2 * h * base), (qp - qm) / (2 * h * base)
            an_l, an_q = getattr(gd, "dlnkh_d" + kind)[t], getattr(gd, "dqst_d" + kind)[t]
            print(f"{TYPES[t][0]} {kind}: dlnKH fd {fd_l:.8g} an {an_l:.8g}; dqst fd {fd_q:.8g} an {an_q:.8g}", flush=True)
            assert abs(an_l - fd_l) <= RTOL * abs(fd_l) + 1e-9
            assert abs(an_q - fd_q) <= RTOL * abs(fd_q) + 1e-9
    # one unit-cell charge, all its images moving together
    nu, a = unit_cell.natoms, 3
    res = []
    for d in (h, -h):
        q = np.array(unit_cell.charges)
        q[a] += d
        uc = pa.Framework(unit_cell.cell, unit_cell.frac, unit_cell.labels, unit_cell.symbols, q)
        res.append(_lnk_qst_dev(pa.replicate(uc, (2, 2, 2)), ff, ns, seed))
    fd_l = (res[0][0] - res[1][0]) / (2 * h)
    fd_q = (res[0][1] - res[1][1]) / (2 * h)
    print(f"atom {a}: dlnKH fd {fd_l:.8g} an {8 * gd.dlnkh_dq[a]:.8g}; dqst fd {fd_q:.8g} an {8 * gd.dqst_dq[a]:.8g}")
    assert abs(8 * gd.dlnkh_dq[a] - fd_l) <= RTOL * abs(fd_l)
    assert abs(8 * gd.dqst_dq[a] - fd_q) <= RTOL * abs(fd_q) + 1e-6


def test_grad_folded_and_generic_reciprocal_sums_agree(unit_cell):
    """CO2 listed C, O, O takes the folded guest structure factor e^{i t0}(q0 + 2 q1 cos p) in the
    gradient's k-space sums; the same molecule listed O, C, O takes the per-site sum. Same poses
    and weights up to FP32 rounding of the energies, so the gradients agree ten times more closely
    than the 1e-3 bar, and both agree with the oracle."""
    fw = pa.replicate(unit_cell, (2, 2, 2))
    ff = _ff()
    g = synth.co2_guest(ff)
    perm = [1, 0, 2]
    ocq = pa.Guest(np.asarray(g.sites)[perm], np.asarray(g.types)[perm], np.asarray(g.charges)[perm])
    ew = pa.EwaldParams(11.0)
    bf = pa.FrameworkBatch([fw], ff, g, ew)
    bg = pa.FrameworkBatch([fw], ff, ocq, ew)
    bf.bin_scale = bg.bin_scale = 1.0
    ns = 20000
    gf = pa.widom_grad(bf, g, T=T, ninsert=ns, seed=3, nblocks=4, precision="mixed")[0]
    gg = pa.widom_grad(bg, ocq, T=T, ninsert=ns, seed=3, nblocks=4, precision="mixed")[0]
    ref = oracle_for(fw, ff, g, ew).widom_grad(T, ns, 3, 0, nthreads=oracle_threads())
    for key in ("dlnkh_dq", "dqst_dq"):
        a, c = getattr(gf, key), getattr(gg, key)
        d = np.abs(a - c).max() / np.abs(c).max()
        print(f"{key}: folded vs generic max |d| / max |entry| = {d:.2e}", flush=True)
        assert d <= 1e-4, key
        _close_atoms(a, ref[key], f"folded {key}")
        _close_atoms(c, ref[key], f"generic {key}")
    for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
        a, c = getattr(gf, key), getattr(gg, key)
        assert np.all(np.abs(a - c) <= 1e-4 * np.abs(c) + 1e-9), key


def test_grad_batch_of_frameworks_against_oracle():
    """Four config-2-style frameworks in one batch, system-major charge layout."""
    fws = synth.config2_frameworks(256)
    pick = [3, 40, 77, 200]
    sub = [fws[i] for i in pick]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0)
    b = pa.FrameworkBatch(sub, ff, g, ew)
    ninsert = 4 * 3000
    gd = pa.widom_grad(b, g, T=T, ninsert=ninsert, seed=21, nblocks=4, precision="mixed")
    ns = pa.system_counts(ninsert, len(sub), pa.default_run(ninsert, len(sub)))
    for s, fw in enumerate(sub):
        ref = oracle_for(fw, ff, g, ew).widom_grad(T, int(ns[s]), 21, s, nthreads=oracle_threads())
        print(f"fw {pick[s]}: {fw.natoms} atoms, ln K_H device {gd[s].result.ln_K_H:.6f} oracle {ref['ln_K_H']:.6f}",
              flush=True)
        if ref["ln_K_H"] < -50:
            continue  # nothing binds: the gradient is that of one repulsive pose
        for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
            _close_types(getattr(gd[s], key), ref[key], f"fw {pick[s]} {key}")
        for key in ("dlnkh_dq", "dqst_dq"):
            _close_atoms(getattr(gd[s], key), ref[key], f"fw {pick[s]} {key}")


def test_grad_uncharged_guest_with_tail(config0):
    fw, _, _, ew = config0
    ff = synth.synthetic_forcefield(12.8, guest="ch4", tail=True)
    g = synth.ch4_guest(ff)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    ns = 20000
    gd = pa.widom_grad(b, g, T=T, ninsert=ns, seed=4, nblocks=4, precision="mixed")[0]
    ref = oracle_for(fw, ff, g, ew).widom_grad(T, ns, 4, 0, nthreads=oracle_threads())
    for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
        _close_types(getattr(gd, key), ref[key], key)
    # an uncharged guest sees no framework charge
    assert np.all(gd.dlnkh_dq == 0.0) and np.all(ref["dlnkh_dq"] == 0.0)


def _ff_scaled(ff, t, kind, factor):
    """`ff` with type t's own epsilon or sigma scaled by `factor`, Lorentz-Berthelot re-mixed."""
    eps = np.diag(ff.epsilon).copy()
    sig = np.diag(ff.sigma).copy()
    (eps if kind == "eps" else sig)[t] *= factor
    return pa.ForceField.from_types(ff.names, list(sig), list(eps), cutoff=ff.cutoff, tail=ff.tail)


def test_grad_config3_binding_frameworks_against_oracle_and_fd():
    """Config 3's frameworks (the first 64 of the 256-framework draw, CO2, LJ + Ewald): the four
    smallest whose device ln K_H shows binding insertions (> -50) at 20,480 insertions each, against
    the FP64 oracle gradient; on the smallest of them, every framework type's epsilon and sigma
    and one unit-cell charge (all its images) against central finite differences of the device's
    FP64 Widom (relative step 1e-4, common random numbers)."""
    fws = synth.config2_frameworks(256)[:64]
    ff = synth.synthetic_forcefield(12.0)
    g = synth.co2_guest(ff)
    ew = pa.EwaldParams(12.0)
    per = 20_480  # a multiple of the 256-insertion run: every system gets exactly this many
    scan = pa.widom(pa.FrameworkBatch(fws, ff, g, ew), g, T=T, ninsert=64 * per, seed=21, nblocks=4, precision="mixed")
    binding = sorted((fws[i].natoms, i) for i in range(64) if scan[i].ln_K_H > -50)
    print(f"config 3: {len(binding)} of 64 frameworks bind; smallest {binding[:4]}", flush=True)
    assert len(binding) >= 4
    pick = [i for _, i in binding[:4]]
    sub = [fws[i] for i in pick]
    b = pa.FrameworkBatch(sub, ff, g, ew)
    gd = pa.widom_grad(b, g, T=T, ninsert=4 * per, seed=21, nblocks=4, precision="mixed")
    ns = pa.system_counts(4 * per, 4, pa.default_run(4 * per, 4))
    for s, fw in enumerate(sub):
        assert ns[s] == per
        ref = oracle_for(fw, ff, g, ew).widom_grad(T, int(ns[s]), 21, s, nthreads=oracle_threads())
        print(f"fw {pick[s]}: {fw.natoms} atoms, {ns[s]} insertions, ln K_H device {gd[s].result.ln_K_H:.6f} oracle "
              f"{ref['ln_K_H']:.6f}", flush=True)
        assert ref["ln_K_H"] > -50
        assert abs(gd[s].result.ln_K_H - ref["ln_K_H"]) <= 1e-4
        for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
            _close_types(getattr(gd[s], key), ref[key], f"fw {pick[s]} {key}")
        for key in ("dlnkh_dq", "dqst_dq"):
            _close_atoms(getattr(gd[s], key), ref[key], f"fw {pick[s]} {key}")

    # finite differences on the smallest binding framework, FP64 device path throughout
    fw = sub[0]
    h, seed = 1e-4, 23

    def lnk_qst(ffx, fwx):
        bx = pa.FrameworkBatch([fwx], ffx, synth.co2_guest(ffx), ew)
        r = pa.widom(bx, synth.co2_guest(ffx), T=T, ninsert=per, seed=seed, nblocks=4, precision="f64")[0]
        return r.ln_K_H, r.q_st

    g1 = pa.widom_grad(pa.FrameworkBatch([fw], ff, g, ew), g, T=T, ninsert=per, seed=seed, nblocks=4,
                       precision="f64")[0]
    for t in range(4):
        for kind in ("eps", "sig"):
            lp, qp = lnk_qst(_ff_scaled(ff, t, kind, 1 + h), fw)
            lm, qm = lnk_qst(_ff_scaled(ff, t, kind, 1 - h), fw)
            base = ff.epsilon[t, t] if kind == "eps" else ff.sigma[t, t]
            fd_l, fd_q = (lp - lm) / (2 * h * base), (qp - qm) / (2 * h * base)
            an_l, an_q = getattr(g1, "dlnkh_d" + kind)[t], getattr(g1, "dqst_d" + kind)[t]
            print(f"{ff.names[t]} {kind}: dlnKH fd {fd_l:.8g} an {an_l:.8g}; dqst fd {fd_q:.8g} an {an_q:.8g}", flush=True)
            assert abs(an_l - fd_l) <= RTOL * abs(fd_l) + 1e-9
            assert abs(an_q - fd_q) <= RTOL * abs(fd_q) + 1e-9
    nimg = int(np.prod(fw.replication))
    nu = fw.natoms // nimg
    a = int(np.argmax(np.abs(g1.dlnkh_dq[:nu])))
    res = []
    for d in (h, -h):
        q = np.array(fw.charges)
        q[a::nu] += d
        res.append(lnk_qst(ff, pa.Framework(fw.cell, fw.frac, fw.labels, fw.symbols, q, fw.replication)))
    fd_l = (res[0][0] - res[1][0]) / (2 * h)
    fd_q = (res[0][1] - res[1][1]) / (2 * h)
    print(f"atom {a} ({nimg} images): dlnKH fd {fd_l:.8g} an {nimg * g1.dlnkh_dq[a]:.8g}; dqst fd {fd_q:.8g} an "
          f"{nimg * g1.dqst_dq[a]:.8g}", flush=True)
    assert abs(nimg * g1.dlnkh_dq[a] - fd_l) <= RTOL * abs(fd_l)
    assert abs(nimg * g1.dqst_dq[a] - fd_q) <= RTOL * abs(fd_q) + 1e-6
