"""Pins the oracle's reweighting gradients (or_widom_grad) by central finite differences of the
oracle's own ln K_H and q_st on the same insertions (common random numbers): perturbed per-type
LJ parameters go through Lorentz-Berthelot mixing again, perturbed charges move every
replication image of a unit-cell atom, and the finite difference of that tied move equals the
image count times the gradient entry."""
import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from pureadsorb_amd import synth

from conftest import oracle_for, oracle_threads

T = 298.15
NS = 3000
SEED = 11
TYPES = [("C_", 52.8, 3.43), ("H_", 22.1, 2.57), ("O_", 30.2, 3.12), ("Zn_", 62.4, 2.46), ("C_co2", 27.0, 2.80),
         ("O_co2", 79.0, 3.05)]


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


def _lnk_qst(fw, ff):
    g = synth.co2_guest(ff)
    o = oracle_for(fw, ff, g, pa.EwaldParams(11.0))
    st, out, _ = o.widom(T, NS, SEED, 0, 4, nthreads=oracle_threads())
    assert st == 0
    return np.log(out[2]), out[4]


def _richardson(at, h, base):
    """Central differences at steps h and h/2 (relative to `base`), Richardson-extrapolated:
    the O(h^2) truncation error cancels, leaving O(h^4)."""
    out = []
    for step in (h, h / 2):
        lp, qp = at(step)
        lm, qm = at(-step)
        out.append(((lp - lm) / (2 * step * base), (qp - qm) / (2 * step * base)))
    return tuple((4 * b - a) / 3 for a, b in zip(out[0], out[1]))


def _grad(fw, ff):
    g = synth.co2_guest(ff)
    o = oracle_for(fw, ff, g, pa.EwaldParams(11.0))
    return o.widom_grad(T, NS, SEED, 0, nthreads=oracle_threads())


@pytest.mark.parametrize("tail", [False, True])
def test_lj_gradients_match_finite_differences(unit_cell, tail):
    fw = pa.replicate(unit_cell, (2, 2, 2))
    ff = _ff(tail=tail)
    gr = _grad(fw, ff)
    lnk0, q0 = _lnk_qst(fw, ff)
    assert abs(gr["ln_K_H"] - lnk0) < 1e-9 * abs(lnk0) and abs(gr["q_st"] - q0) < 1e-12
    for t, (name, e, s) in enumerate(TYPES):
        for kind in ("eps", "sig"):
            def at(h):
                sc = np.ones(len(TYPES))
                sc[t] += h
                return _lnk_qst(fw, _ff(tail=tail, **({"eps_scale": sc} if kind == "eps" else {"sig_scale": sc})))
            base = e * pa.KB if kind == "eps" else s
            fd_l, fd_q = _richardson(at, 1e-4, base)
            an_l = gr["dlnkh_d" + kind][t]
            an_q = gr["dqst_d" + kind][t]
            print(f"{name} {kind}: dlnKH fd {fd_l:.10g} an {an_l:.10g}; dqst fd {fd_q:.10g} an {an_q:.10g}")
            assert abs(an_l - fd_l) <= 1e-7 * abs(fd_l) + 1e-8 * abs(lnk0) / base
            assert abs(an_q - fd_q) <= 1e-7 * abs(fd_q) + 1e-11 / base


def test_charge_gradients_match_tied_finite_differences(unit_cell):
    rep = (2, 2, 2)
    nimg = 8
    fw = pa.replicate(unit_cell, rep)
    ff = _ff()
    gr = _grad(fw, ff)
    nu = unit_cell.natoms
    # every image carries the same entry
    for key in ("dlnkh_dq", "dqst_dq"):
        v = gr[key].reshape(nimg, nu)
        assert np.allclose(v, v[0], rtol=1e-12, atol=1e-14 * np.abs(v).max())
    scale = np.abs(gr["dlnkh_dq"]).max()
    qscale = np.abs(gr["dqst_dq"]).max()
    for a in (0, 7, 19, nu - 1):
        def at(d):
            q = np.array(unit_cell.charges, dtype=np.float64)
            q[a] += d
            uc = pa.Framework(unit_cell.cell, unit_cell.frac, unit_cell.labels, unit_cell.symbols, q)
            return _lnk_qst(pa.replicate(uc, rep), ff)
        fd_l, fd_q = _richardson(at, 1e-4, 1.0)
        print(f"atom {a}: dlnKH fd {fd_l:.10g} an {nimg * gr['dlnkh_dq'][a]:.10g}; dqst fd {fd_q:.10g} "
              f"an {nimg * gr['dqst_dq'][a]:.10g}")
        assert abs(nimg * gr["dlnkh_dq"][a] - fd_l) <= 1e-7 * abs(fd_l) + 1e-9 * scale * nimg
        assert abs(nimg * gr["dqst_dq"][a] - fd_q) <= 1e-7 * abs(fd_q) + 1e-9 * qscale * nimg


def test_replication_mean_equals_unreplicated_direction(unit_cell):
    """With replication (1, 1, 1) (no claim) the charge entries are plain per-atom derivatives;
    the same atoms with a (2, 2, 2) claim report the per-image mean, which the FULL-k oracle
    forms by averaging and which sums back to the tied derivative."""
    fw = pa.replicate(unit_cell, (2, 2, 2))
    ff = _ff()
    g_claim = _grad(fw, ff)
    plain = pa.Framework(fw.cell, fw.frac, fw.labels, fw.symbols, fw.charges)  # replication (1, 1, 1)
    g_plain = _grad(plain, ff)
    v = g_plain["dlnkh_dq"].reshape(8, -1).mean(axis=0)
    assert np.allclose(np.tile(v, 8), g_claim["dlnkh_dq"], rtol=1e-10, atol=1e-12 * np.abs(v).max())
    # LJ gradients do not depend on the claim
    assert np.allclose(g_plain["dlnkh_deps"], g_claim["dlnkh_deps"], rtol=1e-10)
