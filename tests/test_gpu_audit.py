"""The energy audit on the device (audit_energy!, src/audit.jl:97-165) and its fault injection
(after test/audit_tests.jl:60-415).

The device rebuild is held to two independent references: the running structure factor after
an audit equals the framework's plus every guest's sum_s q_s e^{i k.r_s}, formed here in numpy
from the poses and the oracle's k table; the energy after an audit equals pa_total_energy, the
serial FP64 host rebuild. Each corruption goes in through the test-only pa_state_set_value, never
through the device, and each must fail the audit with the reference's message: a corrupted
running energy, a corrupted structure-factor entry, a stale host-energy cache entry (the next
accepted move of that guest folds the wrong dU into the running energy) and an occupancy above
the capacity."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import pureadsorb_amd as pa
from conftest import oracle_for

pytestmark = pytest.mark.gpu

T = 298.15


@pytest.fixture(scope="module")
def rubtak_system(rubtak, trappe, co2):
    sc = pa.replicate(rubtak, (3, 3, 3))
    ew = pa.EwaldParams(12.0, 1e-6)
    b = pa.FrameworkBatch([sc], trappe, co2, ew, fullk=True)
    return sc, trappe, co2, ew, b


def _chain(b, g, n, seed, moves=200, capacities=None):
    st = pa.SystemState(b, g, [n], T=T, seed=seed, capacities=capacities)
    if moves:
        pa.mc_moves(st, pa.MOVE_TRANSLATION, moves, T=T, step_trans=0.5)
        pa.mc_moves(st, pa.MOVE_ROTATION, moves, T=T, step_rot=0.3)
    return st


def test_device_rebuild_against_host_and_poses(rubtak_system):
    sc, ff, g, ew, b = rubtak_system
    st = _chain(b, g, 6, 9)
    e_host = pa.total_energy(st)[0]
    d = pa.audit_energy(st)[0]
    e_dev = st.energy()[0]
    print(f"6 CO2 after 400 moves: device total {e_dev:.15f} eV, host {e_host:.15f} eV, audit discrepancy {d:.2e} eV",
          flush=True)
    assert abs(e_dev - e_host) <= 1e-11 * (1.0 + abs(e_host))
    assert abs(d) < 1e-9
    # the structure factor the audit rebuilt, against numpy from the poses
    o = oracle_for(sc, ff, g, ew)
    kv, _, S, _ = o.ktable(full=True)
    p, q, _ = st.guests(0)
    sites = np.asarray(g.sites, dtype=np.float64)
    rs = np.array([pi + O.rotate(qi, s) for pi, qi in zip(p, q) for s in sites])
    qs = np.tile(np.asarray(g.charges, dtype=np.float64), len(p))
    ref = S + np.exp(1j * (rs @ kv.T)).T @ qs
    sk = st.sk(0)
    err = np.max(np.abs(sk - ref)) / (np.sum(np.abs(sc.charges)) + np.sum(np.abs(qs)))
    print(f"rebuilt structure factor ({len(S)} k-vectors) vs numpy: max |dS| / sum|q| = {err:.2e}", flush=True)
    assert err <= 1e-12


def test_corrupted_running_energy_fails_the_audit(rubtak_system):
    *_, g, _, b = rubtak_system
    st = _chain(b, g, 4, 9)
    pa.audit_energy(st)
    e = st.energy()[0]
    pa.set_state_value(st, "energy", 0, e + 1e-6)
    with pytest.raises(pa.ArgumentError, match=r"energy audit failed for system 1 \(Float64\): accumulated=.*"
                                               r"recomputed=.*exceeds tolerance .* accepted move\(s\) since the last audit"):
        pa.audit_energy(st)


def test_corrupted_structure_factor_fails_the_audit(rubtak_system):
    *_, g, _, b = rubtak_system
    st = _chain(b, g, 4, 11)
    pa.audit_energy(st)
    k = 17
    s = st.sk(0)[k]
    pa.set_state_value(st, "sk", 0, s + 1e-3, index=k)
    with pytest.raises(pa.ArgumentError, match=rf"energy audit failed for system 1 \(Float64\): structure factor at "
                                               rf"k-vector {k + 1} disagrees with the poses: running=.*rebuilt=.*"
                                               r"exceeds tolerance"):
        pa.audit_energy(st)


def test_stale_host_energy_cache_fails_the_audit(rubtak_system):
    *_, g, _, b = rubtak_system
    st = _chain(b, g, 1, 21, moves=0)
    pa.audit_energy(st)
    _, _, he = st.guests(0)
    pa.set_state_value(st, "host_energy", 0, he[0] + 1.0, index=0)  # stale by 1 eV
    # the next accepted move of the (only) guest takes dU against the stale cached value
    tr = pa.mc_moves(st, pa.MOVE_TRANSLATION, 20, T=T, step_trans=0.5, trace=True)[0]
    assert tr[:, 3].sum() >= 1
    with pytest.raises(pa.ArgumentError, match=r"energy audit failed for system 1 \(Float64\): accumulated="):
        pa.audit_energy(st)


def test_occupancy_above_capacity_fails_the_audit(rubtak_system):
    *_, g, _, b = rubtak_system
    st = _chain(b, g, 2, 9, moves=0, capacities=[5])
    pa.audit_energy(st)
    pa.set_state_value(st, "occupancy", 0, 6)
    with pytest.raises(pa.ArgumentError, match=r"energy audit failed for system 1 \(Float64\): occupancy 6 exceeds "
                                               r"capacity 5"):
        pa.audit_energy(st)
    # nothing else reads the guest slots of such a state
    with pytest.raises(pa.ArgumentError, match="exceeds its capacity"):
        pa.mc_step(st, pa.MOVE_TRANSLATION, 0.5, 0.3, T=T)
    with pytest.raises(pa.ArgumentError, match="exceeds its capacity"):
        pa.total_energy(st)
    # raising it within the capacity would expose unset slots
    with pytest.raises(pa.ArgumentError, match="unset guest slots"):
        pa.set_state_value(st, "occupancy", 0, 4)
    pa.set_state_value(st, "occupancy", 0, 2)
    assert abs(pa.audit_energy(st)[0]) < 1e-9
    pa.mc_step(st, pa.MOVE_TRANSLATION, 0.5, 0.3, T=T)
