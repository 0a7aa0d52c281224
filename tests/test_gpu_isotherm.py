"""GPU run of the isotherm driver (src/isotherm.jl:104-150): the ideal-gas law at every pressure
(test/isotherm_tests.jl:75-108: <N> = f V / kT with f from the equation of state, phi ~ 1 for a
guest with pc = 1e12 Pa) and the capacity failure propagated from the exchange moves."""
from __future__ import annotations

import numpy as np
import pytest

import pureadsorb_amd as pa

pytestmark = pytest.mark.gpu


def _ideal_box(V):
    L = V ** (1 / 3)
    fw = pa.Framework(np.diag([L, L, L]), [[0.5, 0.5, 0.5]], ["X1"], ["X"], [0.0])
    ff = pa.ForceField.from_types(["X_"], [1.0], [0.0], cutoff=8.0, tail=False)
    g = pa.Guest([[0.0, 0.0, 0.0]], [0], [0.0], 300.0, 1.0e12, 0.0)
    return fw, ff, g


def test_isotherm_reproduces_the_ideal_gas_law():
    V, T = 37219.0, 273.15
    fw, ff, g = _ideal_box(V)
    pressures = [5.0e4, 1.01325e5, 2.0e5]
    iso = pa.run_isotherm(fw, ff, g, pa.EwaldParams(8.0), T=T, pressures=pressures, nreplicas=400, capacity=40,
                          n_warmup=200, n_production=500, n_audit=701, step_trans=0.3, step_rot=0.3,
                          exchange_prob=1.0, seed=7, min_cycle_length=30)
    assert list(iso.pressure) == pressures and iso.nreplicas == 400
    kT = pa.KB * T
    for p in range(3):
        target = pressures[p] * pa.PASCAL * V / kT
        print(f"P {pressures[p]:.4g} Pa: <N> {iso.loading[p]:.4f} +- {iso.loading_err[p]:.4f}, ideal {target:.4f}",
              flush=True)
        assert abs(iso.loading[p] - target) <= max(8 * iso.loading_err[p], 0.15 * target)
        assert iso.max_occupancy[p] <= iso.capacity[p]
    assert np.all(np.diff(iso.loading) > 0)


def test_isotherm_propagates_the_capacity_failure():
    fw, ff, g = _ideal_box(8000.0)
    with pytest.raises(pa.ArgumentError, match="hit capacity"):
        pa.run_isotherm(fw, ff, g, pa.EwaldParams(8.0), T=298.15, pressures=[1.0e9], nreplicas=1, capacity=3,
                        n_warmup=0, n_production=50, n_audit=1000, step_trans=0.3, step_rot=0.3, exchange_prob=1.0,
                        seed=4, nblocks=2)
