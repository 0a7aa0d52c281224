"""MI355X (gfx950) backend of PureAdsorb.jl's Widom insertion path.

The public names mirror PureAdsorb.jl's API (src/PureAdsorb.jl:27-37): `Framework`,
`read_cif`, `replicate`, `ForceField`, `read_forcefield`, `Guest`, `read_guest`,
`EwaldParams`, `FrameworkBatch`, `widom`, `WidomResult`. The compute runs in
libpureadsorb_amd.so (HIP kernels for gfx950) through its C ABI; there is no CPU fallback.
"""
from ._lib import ArgumentError, DeviceError, DimensionMismatch, device_name, synchronize
from .batch import FrameworkBatch, nframeworks
from .forcefield import EwaldParams, ForceField, Guest, read_forcefield, read_guest, tail_coefficient, tail_delta, typeindex
from .fugacity import Fugacity, cubic_roots, peng_robinson_fugacity
from .isotherm import IsothermResult, combine_replicas, run_isotherm
from .io_native import cif_types_native, read_cif_native, read_forcefield_native, read_guest_native
from .mc import (MOVE_DELETION, MOVE_INSERTION, MOVE_REINSERTION, MOVE_ROTATION, MOVE_TRANSLATION, GCMCResult,
                 NVTResult, SystemState, audit_energy, gcmc_samples, mc_delete, mc_exchange, mc_insert, mc_moves, mc_step,
                 nvt_last_timing, nvt_samples, run_gcmc, run_nvt, set_state_value, total_energy, widom_chain_energies)
from .structure import (Framework, cartesian, cell_matrix, min_multiplicity, natoms, perpendicular_lengths, read_cif,
                        reciprocal_basis, replicate, total_charge, volume)
from .widom import (KB, KE, PASCAL, WidomGrad, WidomResult, default_run, hardcore_flags, last_timing, sys_of_index,
                    system_counts, widom, widom_energies, widom_grad, widom_poses)

__all__ = [
    "ArgumentError", "DeviceError", "DimensionMismatch", "device_name", "synchronize", "FrameworkBatch", "nframeworks", "EwaldParams",
    "ForceField", "Guest", "read_forcefield", "read_guest", "tail_coefficient", "tail_delta", "typeindex", "Framework",
    "cartesian", "cell_matrix", "min_multiplicity", "natoms", "perpendicular_lengths", "read_cif", "reciprocal_basis",
    "replicate", "total_charge", "volume", "KB", "KE", "PASCAL", "WidomResult", "default_run", "last_timing",
    "sys_of_index", "system_counts", "widom", "widom_energies", "widom_grad", "widom_poses", "WidomGrad", "hardcore_flags",
    "MOVE_TRANSLATION", "MOVE_ROTATION", "MOVE_REINSERTION", "MOVE_INSERTION", "MOVE_DELETION", "SystemState",
    "mc_step", "mc_insert", "mc_delete", "mc_exchange", "mc_moves", "total_energy", "audit_energy", "set_state_value", "run_gcmc",
    "GCMCResult", "gcmc_samples", "NVTResult", "run_nvt", "nvt_samples", "widom_chain_energies", "nvt_last_timing",
    "Fugacity", "cubic_roots", "peng_robinson_fugacity", "IsothermResult", "combine_replicas", "run_isotherm",
    "read_cif_native", "read_forcefield_native", "read_guest_native", "cif_types_native",
]
