"""GPU parity for guests of every supported site count (the kernels are instantiated for 1 to 6
sites): FP64 and mixed insertion energies against the oracle, phase-0 flags, the mixed Widom
result against the FP64 path, move traces of the FP64 move kernel against the oracle chain, and
test insertions into the loaded chain plus a short NVT run against the oracle's.

Bars as in test_gpu_widom.py and test_gpu_mc.py: FP64 energies within 1e-11 relative; mixed
energies within 1e-5 relative below 1e3 eV (3e-4 on sub-angstrom contacts); every insertion the
device rejects in phase 0 is rejected by the oracle and has weight exactly 0.0; move slots and
decisions equal, dU within 1e-9 relative."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from conftest import oracle_for, oracle_threads
from pureadsorb_amd import synth

pytestmark = pytest.mark.gpu

T = 298.15
KT = pa.KB * T

# site positions (A), type names, charges (e): a charged ion, an uncharged dumbbell, and charge
# distributions without a dipole (a planar triangle, a tetrahedron, an octahedron), whose
# orientation-averaged Ewald self term stays within the reference's tolerance
_TET = 1.09 / np.sqrt(3.0)
GUESTS = {
    1: ([[0.0, 0.0, 0.0]], ["C_co2"], [0.3]),
    2: ([[-0.6, 0.0, 0.0], [0.6, 0.0, 0.0]], ["O_co2", "O_co2"], [0.0, 0.0]),
    4: ([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [-0.5, 0.8660254, 0.0], [-0.5, -0.8660254, 0.0]],
        ["C_co2", "O_co2", "O_co2", "O_co2"], [0.6, -0.2, -0.2, -0.2]),
    5: ([[0.0, 0.0, 0.0], [_TET, _TET, _TET], [_TET, -_TET, -_TET], [-_TET, _TET, -_TET], [-_TET, -_TET, _TET]],
        ["C_co2", "O_co2", "O_co2", "O_co2", "O_co2"], [0.8, -0.2, -0.2, -0.2, -0.2]),
    6: ([[1.1, 0.0, 0.0], [-1.1, 0.0, 0.0], [0.0, 1.1, 0.0], [0.0, -1.1, 0.0], [0.0, 0.0, 1.1], [0.0, 0.0, -1.1]],
        ["O_co2", "O_co2", "C_co2", "C_co2", "C_co2", "C_co2"], [-0.12, -0.12, 0.06, 0.06, 0.06, 0.06]),
}


def _guest(ff, ns):
    sites, names, q = GUESTS[ns]
    return pa.Guest(sites, [ff.names.index(n) for n in names], q)


@pytest.fixture(scope="module")
def porous_cell():
    ff = synth.synthetic_forcefield(12.0)
    fw = pa.replicate(synth.synthetic_framework(13.0, 13.0, 15.0, 90, 90, 120, 80, 42, charge_sd=0.5), (3, 3, 3))
    return fw, ff, pa.EwaldParams(12.0, 1e-6)


@pytest.mark.parametrize("ns", sorted(GUESTS))
def test_guest_energies_and_widom(ns, porous_cell):
    fw, ff, ew = porous_cell
    g = _guest(ff, ns)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    o = oracle_for(fw, ff, g, ew)
    n = 1500
    ref = o.energies(13, 0, 0, n, nthreads=oracle_threads())
    d64 = pa.widom_energies(b, 0, 13, 0, n, precision="f64")
    dmx = pa.widom_energies(b, 0, 13, 0, n, precision="mixed")
    fin = np.isfinite(ref)
    rel64 = np.abs(d64[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1.0)
    relmx = np.abs(dmx[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1.0)
    mod = np.abs(ref[fin]) < 1e3
    print(f"{ns} sites: f64 max rel {rel64.max():.2e}, mixed {relmx[mod].max():.2e} below 1e3 eV "
          f"({mod.sum()} poses), {relmx.max():.2e} overall", flush=True)
    assert np.array_equal(np.isfinite(d64), fin) and np.array_equal(np.isfinite(dmx), fin)
    assert rel64.max() < 1e-11
    assert relmx[mod].max() < 1e-5 and relmx.max() < 3e-4
    # phase 0: the device's rejections are the oracle's, with weight exactly zero
    flags = pa.hardcore_flags(b, 0, 13, T, 0, n).astype(bool)
    _, rho2 = o.rejection_tables(KT, 1.1920929e-07)
    oflags = o.hardcore_flags(13, 0, 0, n, rho2, nthreads=oracle_threads()).astype(bool)
    assert not np.any(flags & ~oflags)
    assert np.all(np.exp(-ref[flags] / KT) == 0.0)
    # the two-phase mixed Widom result against the FP64 path on the same streams
    rm = pa.widom(b, g, T=T, ninsert=60_000, seed=21, nblocks=4, precision="mixed")[0]
    r64 = pa.widom(b, g, T=T, ninsert=60_000, seed=21, nblocks=4, precision="f64")[0]
    print(f"{ns} sites: ln K_H mixed {rm.ln_K_H:.6f} f64 {r64.ln_K_H:.6f}", flush=True)
    assert abs(rm.ln_K_H - r64.ln_K_H) <= 1e-4 * max(1.0, abs(r64.ln_K_H) / 100.0)


@pytest.mark.parametrize("ns", sorted(GUESTS))
def test_guest_move_traces(ns):
    ff = synth.synthetic_forcefield(11.0)
    g = _guest(ff, ns)
    uc = synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5)
    fw = pa.replicate(uc, (2, 2, 2))
    ew = pa.EwaldParams(11.0)
    b = pa.FrameworkBatch([fw], ff, g, ew, fullk=True)
    st = pa.SystemState(b, g, [4], T=T, seed=9, capacities=[30])
    o = oracle_for(fw, ff, g, ew)
    p, q, _ = st.guests(0)
    m = O.OracleMC(o, 30, O.splitmix64(9, 1), p, q)
    for mt, f in ((1, 1e5), (2, 1e5), (4, 1e8), (3, 1e5), (5, 1e2)):
        tr = pa.mc_moves(st, mt, 24, T=T, step_trans=0.3, step_rot=0.2, fugacity=f, trace=True)[0]
        ref = np.array([m.attempt(mt, 0.3, 0.2, f, T) for _ in range(24)])
        assert np.array_equal(tr[:, 3].astype(int), ref[:, 0].astype(int)), (ns, mt)
        fin = np.isfinite(ref[:, 1])
        assert np.array_equal(np.isfinite(tr[:, 2]), fin)
        assert np.all(np.abs(tr[fin, 2] - ref[fin, 1]) <= 1e-9 * (1 + np.abs(ref[fin, 1]))), (ns, mt)
    info = st.info(0)
    print(f"{ns} sites: {info['occupancy']} guests, accepted {info['accepted']}", flush=True)
    assert info["occupancy"] == m.scalars()["occupancy"]
    assert abs(info["energy"] - m.scalars()["energy"]) <= 1e-9 * (1 + abs(info["energy"]))
    if np.any(np.asarray(g.charges) != 0.0):
        # an uncharged guest keeps no structure factor on the device
        assert np.allclose(st.sk(0), m.sk(), rtol=0, atol=1e-9)
    assert np.all(np.abs(pa.audit_energy(st)) < 1e-8)
    # test insertions into the loaded configuration and a short NVT run, against the oracle chain
    rng = np.random.default_rng(ns)
    pos = rng.random((100, 3)) @ o.cell.T
    qq = rng.normal(size=(100, 4))
    qq /= np.linalg.norm(qq, axis=1)[:, None]
    dU = pa.widom_chain_energies(st, 0, pos, qq)
    refc = np.array([m.chain_insertion(x, y) for x, y in zip(pos, qq)])
    fin = np.isfinite(refc) & (np.abs(refc) < 1e8)
    assert fin.sum() > 50
    assert np.all(np.abs(dU[fin] - refc[fin]) <= 1e-9 * (1 + np.abs(refc[fin]))), ns
    kw = dict(T=T, n_warmup=1, n_production=4, n_widom_per_cycle=8, n_audit=2, step_trans=0.3, step_rot=0.2,
              widom_seed=5, nblocks=2)
    r = pa.run_nvt(b, st, g, ff, **kw)[0]
    rf = O.nvt_run([m], **kw)[0]
    _, sw, _ = pa.nvt_samples(st, 0, 4)
    print(f"{ns} sites: NVT mu_ex {r.mu_ex:.9f} (oracle {rf['mu_ex']:.9f}), acceptance {r.acceptance}", flush=True)
    assert r.acceptance == rf["acceptance"]
    assert np.allclose(sw, rf["sW"], rtol=1e-8, atol=1e-300)


@pytest.mark.parametrize("ns", [1, 2, 5])
def test_guest_gradients(ns):
    """pa_widom_grad for other site counts against the oracle's FP64 gradient (1e-3 relative per
    type entry; charge entries within 1e-3 of their largest)."""
    ff = synth.synthetic_forcefield(11.0)
    g = _guest(ff, ns)
    fw = pa.replicate(synth.synthetic_framework(12.0, 12.0, 12.0, 90, 90, 90, 40, 5, charge_sd=0.5), (2, 2, 2))
    ew = pa.EwaldParams(11.0)
    b = pa.FrameworkBatch([fw], ff, g, ew)
    n = 20000
    gd = pa.widom_grad(b, g, T=T, ninsert=n, seed=3, nblocks=4, precision="mixed")[0]
    ref = oracle_for(fw, ff, g, ew).widom_grad(T, n, 3, 0, nthreads=oracle_threads())
    assert abs(gd.result.ln_K_H - ref["ln_K_H"]) <= 1e-4
    for key in ("dlnkh_deps", "dlnkh_dsig", "dqst_deps", "dqst_dsig"):
        an, rf = getattr(gd, key), ref[key]
        print(f"{ns} sites {key}: max rel {np.max(np.abs(an - rf) / np.maximum(np.abs(rf), 1e-9)):.2e}", flush=True)
        assert np.all(np.abs(an - rf) <= 1e-3 * np.abs(rf) + 1e-9), key
    for key in ("dlnkh_dq", "dqst_dq"):
        an, rf = getattr(gd, key), ref[key]
        assert np.abs(an - rf).max() <= 1e-3 * np.abs(rf).max() + 1e-12, key
