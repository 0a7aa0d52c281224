"""Adsorption isotherms (src/isotherm.jl): one batch of (pressure, replica) systems on one
framework, one GCMC run over all of them on the device, replicas folded per pressure."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import ArgumentError, DimensionMismatch
from .batch import FrameworkBatch
from .fugacity import peng_robinson_fugacity
from .mc import SystemState, _T, run_gcmc


def combine_replicas(values, errors) -> tuple[float, float]:
    """`combine_replicas` (src/isotherm.jl:30-38): the mean of R independent estimates and
    sqrt(within^2 + between^2), within^2 = sum(err^2)/R^2, between^2 = var(values)/R (0 at R = 1)."""
    v = [float(x) for x in values]
    e = [float(x) for x in errors]
    R = len(v)
    if R != len(e):
        raise DimensionMismatch("combine_replicas: values and errors must have the same length")
    if R < 1:
        raise ArgumentError("combine_replicas: need at least one replica")
    m = sum(v) / R
    within = sum(x * x for x in e) / (R * R)
    between = sum((x - m) ** 2 for x in v) / ((R - 1) * R) if R > 1 else 0.0
    return m, math.sqrt(within + between)


@dataclass(frozen=True)
class IsothermResult:
    """`IsothermResult` (src/isotherm.jl:55-64): per pressure point (Pa) the replica-combined
    loading and energy (eV) with their errors, the largest occupancy any replica reached and the
    shared capacity."""
    pressure: np.ndarray
    loading: np.ndarray
    loading_err: np.ndarray
    energy: np.ndarray
    energy_err: np.ndarray
    max_occupancy: np.ndarray
    capacity: np.ndarray
    nreplicas: int


def run_isotherm(fw, ff, guest, ewald, *, T, pressures, nreplicas: int, capacity: int, n_warmup: int,
                 n_production: int, n_audit: int, step_trans: float, step_rot: float, exchange_prob: float = 0.5,
                 min_cycle_length: int = 1, seed: int = 0, nblocks: int = 10) -> IsothermResult:
    """`run_isotherm!(fw, ff, guest, ewald; T, pressures, nreplicas, capacity, n_warmup,
    n_production, n_audit, step_trans, step_rot, exchange_prob, min_cycle_length, seed, nblocks)`
    (src/isotherm.jl:104-150): system p * nreplicas + r runs replica r at the Peng-Robinson
    fugacity of pressures[p]; the batch stores the framework once; every chain starts empty with
    `capacity` slots; one run_gcmc over the batch, exchange with probability exchange_prob and
    otherwise translation, rotation or reinsertion uniformly."""
    nreplicas = int(nreplicas)
    if nreplicas < 1:
        raise ArgumentError(f"run_isotherm!: nreplicas={nreplicas} must be >= 1")
    pressures = [float(_T(p)) for p in pressures]
    npress = len(pressures)
    if npress < 1:
        raise ArgumentError("run_isotherm!: pressures must be non-empty")
    nsys = npress * nreplicas
    Tk = _T(T)
    batch = FrameworkBatch([fw] * nsys, ff, guest, ewald, fullk=True)
    fug = np.empty(nsys)
    for p in range(npress):
        fug[p * nreplicas:(p + 1) * nreplicas] = peng_robinson_fugacity(pressures[p], Tk, guest=guest).f
    state = SystemState(batch, guest, [0] * nsys, T=Tk, seed=seed, capacities=[int(capacity)] * nsys)
    pm = (1.0 - float(exchange_prob)) / 3.0
    res = run_gcmc(state, T=Tk, n_warmup=n_warmup, n_production=n_production, n_audit=n_audit,
                   step_trans=float(step_trans), step_rot=float(step_rot), fugacity=fug, p_translation=pm,
                   p_rotation=pm, p_reinsertion=pm, p_exchange=float(exchange_prob),
                   min_cycle_length=min_cycle_length, nblocks=nblocks)
    out = {k: np.empty(npress) for k in ("loading", "loading_err", "energy", "energy_err")}
    mo, cap = np.empty(npress, dtype=int), np.empty(npress, dtype=int)
    for p in range(npress):
        rr = res[p * nreplicas:(p + 1) * nreplicas]
        out["loading"][p], out["loading_err"][p] = combine_replicas([r.loading for r in rr], [r.loading_err for r in rr])
        out["energy"][p], out["energy_err"][p] = combine_replicas([r.energy for r in rr], [r.energy_err for r in rr])
        mo[p] = max(r.max_occupancy for r in rr)
        cap[p] = rr[0].capacity
    state.close()
    batch.close()
    return IsothermResult(np.array(pressures), out["loading"], out["loading_err"], out["energy"], out["energy_err"],
                          mo, cap, nreplicas)
