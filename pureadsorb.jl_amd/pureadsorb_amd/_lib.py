"""ctypes binding of libpureadsorb_amd.so (the C ABI in include/pureadsorb_amd.h).

The library is loaded from this directory. A missing library is an error: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpureadsorb_amd.so")

PA_OK = 0
PA_ERR_ARG = 1
PA_ERR_DIM = 2
PA_ERR_DEVICE = 3
PA_ERR_INTERNAL = 4

PA_PREC_F64 = 0
PA_PREC_MIXED = 1
PA_SINGLE_PHASE = 0x100

# pa_state_set_value fields (test-only overwrites of one element of a chain's running state)
PA_FIELD_ENERGY = 1
PA_FIELD_SK = 2
PA_FIELD_HOST_ENERGY = 3
PA_FIELD_OCCUPANCY = 4

PA_BATCH_FULLK = 1
PA_BATCH_HOST_ONLY = 2

# Every symbol include/pureadsorb_amd.h declares.
EXPORTS = (
    "pa_batch_create", "pa_batch_from_tables", "pa_batch_destroy", "pa_batch_nframeworks", "pa_batch_framework_of",
    "pa_batch_framework_info", "pa_widom", "pa_widom_noalloc", "pa_widom_grad", "pa_widom_energies", "pa_hardcore_flags", "pa_widom_poses", "pa_last_timing",
    "pa_last_error", "pa_device_name", "pa_synchronize",
    "pa_state_create", "pa_state_destroy", "pa_mc_step", "pa_mc_insert", "pa_mc_delete", "pa_mc_moves", "pa_state_info",
    "pa_state_guests", "pa_state_sk", "pa_total_energy", "pa_audit_energy", "pa_state_set_value", "pa_gcmc_run", "pa_gcmc_samples", "pa_batch_set_bin_scale", "pa_batch_bin_scale",
    "pa_batch_set_list_capacity", "pa_batch_list_stats",
    "pa_widom_chain_energies", "pa_nvt_run", "pa_nvt_samples", "pa_nvt_last_timing",
    "pa_read_cif", "pa_cif_info", "pa_cif_atoms", "pa_cif_strings", "pa_cif_types", "pa_cif_destroy",
    "pa_read_forcefield", "pa_ff_info", "pa_ff_params", "pa_ff_names", "pa_ff_typeindex", "pa_ff_destroy",
    "pa_read_guest", "pa_guest_info", "pa_guest_sites", "pa_guest_destroy",
)


class ArgumentError(ValueError):
    """The reference's ArgumentError (a bad argument, caught before any launch)."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


class PaFramework(C.Structure):
    _fields_ = [
        ("cell", C.c_double * 9),
        ("natoms", C.c_int32),
        ("frac", C.c_void_p),
        ("types", C.c_void_p),
        ("charges", C.c_void_p),
        ("replication", C.c_int32 * 3),
    ]


class PaForcefield(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int32),
        ("sigma", C.c_void_p),
        ("epsilon", C.c_void_p),
        ("cutoff", C.c_double),
        ("tail", C.c_int32),
    ]


class PaGuest(C.Structure):
    _fields_ = [
        ("nsites", C.c_int32),
        ("sites", C.c_void_p),
        ("types", C.c_void_p),
        ("charges", C.c_void_p),
    ]


class PaEwald(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("precision", C.c_double)]


class PaWidomResult(C.Structure):
    _fields_ = [
        ("mu_ex", C.c_double), ("mu_ex_err", C.c_double),
        ("K_H", C.c_double), ("K_H_err", C.c_double),
        ("q_st", C.c_double), ("q_st_err", C.c_double),
        ("nsamples", C.c_int64), ("nblocks", C.c_int64),
        ("ln_K_H", C.c_double),
    ]


class PaBatchTables(C.Structure):
    """pa_batch_tables: a FrameworkBatch's arrays (src/batch.jl:72-104), 0-based."""
    _fields_ = [
        ("nframeworks", C.c_int32), ("nsys", C.c_int32), ("framework_of", C.c_void_p), ("atom_offsets", C.c_void_p),
        ("positions", C.c_void_p), ("types", C.c_void_p), ("charges", C.c_void_p), ("cells", C.c_void_p),
        ("alphas", C.c_void_p), ("k_offsets", C.c_void_p), ("ks", C.c_void_p), ("kprefactor", C.c_void_p),
        ("Shost", C.c_void_p), ("constant_offset", C.c_void_p), ("self_term_halfrange", C.c_void_p),
        ("ntypes", C.c_int32), ("sigma", C.c_void_p), ("epsilon", C.c_void_p), ("compact_to_orig", C.c_void_p),
        ("bs", C.c_void_p), ("kmin", C.c_void_p), ("cutoff", C.c_double), ("ewald_cutoff", C.c_double),
        ("fullk", C.c_int32),
    ]


class PaFrameworkInfo(C.Structure):
    _fields_ = [
        ("natoms", C.c_int32), ("nk", C.c_int32), ("nk_full", C.c_int32),
        ("volume", C.c_double), ("alpha", C.c_double),
        ("constant_offset", C.c_double), ("self_term_halfrange", C.c_double),
        ("hardcore_bound", C.c_double),
    ]


class PaChainInfo(C.Structure):
    _fields_ = [
        ("occupancy", C.c_int32), ("capacity", C.c_int32), ("energy", C.c_double), ("energy_abs_accum", C.c_double),
        ("rng_counter", C.c_int64), ("accepted", C.c_int32 * 5), ("attempted", C.c_int32 * 5),
        ("const_p", C.c_double), ("const_q", C.c_double),
    ]


class PaGcmcParams(C.Structure):
    _fields_ = [
        ("T", C.c_double), ("n_warmup", C.c_int64), ("n_production", C.c_int64), ("n_audit", C.c_int64),
        ("min_cycle_length", C.c_int64), ("fugacity_pa", C.c_void_p), ("step_trans", C.c_void_p),
        ("step_rot", C.c_void_p), ("p_translation", C.c_double), ("p_rotation", C.c_double),
        ("p_reinsertion", C.c_double), ("p_exchange", C.c_double), ("nblocks", C.c_int32),
    ]


class PaGcmcResult(C.Structure):
    _fields_ = [
        ("loading", C.c_double), ("loading_err", C.c_double), ("energy", C.c_double), ("energy_err", C.c_double),
        ("q_st", C.c_double), ("q_st_err", C.c_double), ("corr_UN", C.c_double), ("max_occupancy", C.c_int32),
        ("capacity", C.c_int32), ("ncycles", C.c_int64), ("accepted", C.c_int64 * 5), ("attempted", C.c_int64 * 5),
    ]


class PaNvtParams(C.Structure):
    _fields_ = [
        ("T", C.c_double), ("n_warmup", C.c_int64), ("n_production", C.c_int64), ("n_widom_per_cycle", C.c_int64),
        ("n_audit", C.c_int64), ("min_cycle_length", C.c_int64), ("step_trans", C.c_void_p), ("step_rot", C.c_void_p),
        ("widom_seed", C.c_uint64), ("nblocks", C.c_int32),
    ]


class PaNvtResult(C.Structure):
    _fields_ = [
        ("energy", C.c_double), ("energy_err", C.c_double), ("mu_ex", C.c_double), ("mu_ex_err", C.c_double),
        ("K_H", C.c_double), ("K_H_err", C.c_double), ("q_st", C.c_double), ("q_st_err", C.c_double),
        ("acceptance", C.c_double * 3), ("ncycles", C.c_int64),
    ]


class PaNvtTiming(C.Structure):
    _fields_ = [
        ("move_ms", C.c_double), ("move_launches", C.c_int64), ("insertion_ms", C.c_double),
        ("insertion_launches", C.c_int64), ("insertions", C.c_int64), ("total_ms", C.c_double),
    ]


class PaTiming(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double), ("energy_ms", C.c_double), ("energy_launches", C.c_int64),
        ("reduce_ms", C.c_double), ("insertions", C.c_int64), ("binning_ms", C.c_double),
        ("grad_ms", C.c_double), ("grad_insertions", C.c_int64),
    ]


_lib = None


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python pureadsorb.jl_amd/build.py` "
            "(there is no non-HIP fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.pa_batch_create.restype = C.c_int
    L.pa_batch_create.argtypes = [C.POINTER(PaFramework), C.c_int32, C.POINTER(PaForcefield), C.POINTER(PaGuest),
                                  C.POINTER(PaEwald), C.c_uint32, C.POINTER(vp)]
    L.pa_batch_from_tables.restype = C.c_int
    L.pa_batch_from_tables.argtypes = [C.POINTER(PaBatchTables), C.POINTER(PaGuest), C.POINTER(vp)]
    L.pa_batch_destroy.restype = None
    L.pa_batch_destroy.argtypes = [vp]
    L.pa_batch_nframeworks.restype = C.c_int
    L.pa_batch_nframeworks.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.pa_batch_framework_of.restype = C.c_int
    L.pa_batch_framework_of.argtypes = [vp, vp]
    L.pa_batch_framework_info.restype = C.c_int
    L.pa_batch_framework_info.argtypes = [vp, C.c_int32, C.POINTER(PaFrameworkInfo)]
    L.pa_widom.restype = C.c_int
    L.pa_widom.argtypes = [vp, C.POINTER(PaGuest), C.c_double, C.c_int64, C.c_uint64, C.c_int32, C.c_int32, C.c_int64,
                           C.c_uint32, vp]
    L.pa_widom_noalloc.restype = C.c_int
    L.pa_widom_noalloc.argtypes = L.pa_widom.argtypes
    L.pa_widom_grad.restype = C.c_int
    L.pa_widom_grad.argtypes = [vp, C.POINTER(PaGuest), C.c_double, C.c_int64, C.c_uint64, C.c_int32, C.c_int32,
                                C.c_uint32, vp, vp, vp, vp, vp, vp, vp]
    L.pa_batch_set_bin_scale.restype = C.c_int
    L.pa_batch_set_bin_scale.argtypes = [vp, C.c_double]
    L.pa_batch_bin_scale.restype = C.c_int
    L.pa_batch_bin_scale.argtypes = [vp, C.POINTER(C.c_double)]
    L.pa_batch_set_list_capacity.restype = C.c_int
    L.pa_batch_set_list_capacity.argtypes = [vp, C.c_int32]
    L.pa_batch_list_stats.restype = C.c_int
    L.pa_batch_list_stats.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.pa_state_create.restype = C.c_int
    L.pa_state_create.argtypes = [vp, C.POINTER(PaGuest), vp, vp, C.c_double, C.c_uint64, C.POINTER(vp)]
    L.pa_state_destroy.restype = None
    L.pa_state_destroy.argtypes = [vp]
    L.pa_mc_step.restype = C.c_int
    L.pa_mc_step.argtypes = [vp, C.c_int32, vp, vp, C.c_double]
    L.pa_mc_insert.restype = C.c_int
    L.pa_mc_insert.argtypes = [vp, vp, C.c_double]
    L.pa_mc_delete.restype = C.c_int
    L.pa_mc_delete.argtypes = [vp, vp, C.c_double]
    L.pa_mc_moves.restype = C.c_int
    L.pa_mc_moves.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp, C.c_double, vp]
    L.pa_state_info.restype = C.c_int
    L.pa_state_info.argtypes = [vp, C.c_int32, C.POINTER(PaChainInfo)]
    L.pa_state_guests.restype = C.c_int
    L.pa_state_guests.argtypes = [vp, C.c_int32, vp, vp, vp]
    L.pa_state_sk.restype = C.c_int
    L.pa_state_sk.argtypes = [vp, C.c_int32, vp]
    L.pa_total_energy.restype = C.c_int
    L.pa_total_energy.argtypes = [vp, vp]
    L.pa_audit_energy.restype = C.c_int
    L.pa_audit_energy.argtypes = [vp, vp]
    L.pa_state_set_value.restype = C.c_int
    L.pa_state_set_value.argtypes = [vp, C.c_int32, C.c_int32, C.c_int64, C.c_double, C.c_double]
    L.pa_gcmc_run.restype = C.c_int
    L.pa_gcmc_run.argtypes = [vp, C.POINTER(PaGcmcParams), vp]
    L.pa_gcmc_samples.restype = C.c_int
    L.pa_gcmc_samples.argtypes = [vp, C.c_int32, C.c_int64, vp, vp]
    L.pa_widom_chain_energies.restype = C.c_int
    L.pa_widom_chain_energies.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp]
    L.pa_nvt_run.restype = C.c_int
    L.pa_nvt_run.argtypes = [vp, C.POINTER(PaGuest), C.POINTER(PaNvtParams), vp]
    L.pa_nvt_samples.restype = C.c_int
    L.pa_nvt_samples.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp]
    L.pa_nvt_last_timing.restype = C.c_int
    L.pa_nvt_last_timing.argtypes = [vp, C.POINTER(PaNvtTiming)]
    L.pa_hardcore_flags.restype = C.c_int
    L.pa_hardcore_flags.argtypes = [vp, C.c_int32, C.c_uint64, C.c_double, C.c_int64, C.c_int64, vp]
    L.pa_widom_energies.restype = C.c_int
    L.pa_widom_energies.argtypes = [vp, C.c_int32, C.c_uint64, C.c_int64, C.c_int64, C.c_uint32, vp]
    L.pa_widom_poses.restype = C.c_int
    L.pa_widom_poses.argtypes = [vp, C.c_int32, C.c_uint64, C.c_int64, C.c_int64, vp, vp]
    L.pa_last_timing.restype = C.c_int
    L.pa_last_timing.argtypes = [vp, C.POINTER(PaTiming)]
    L.pa_last_error.restype = C.c_int
    L.pa_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.pa_device_name.restype = C.c_int
    L.pa_device_name.argtypes = [C.c_char_p, C.c_size_t]
    L.pa_synchronize.restype = C.c_int
    L.pa_synchronize.argtypes = []
    # input files (host only)
    i32p, dp, szp = C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_size_t)
    for name, args in (
        ("pa_read_cif", [C.c_char_p, C.POINTER(vp)]),
        ("pa_cif_info", [vp, i32p, vp]),
        ("pa_cif_atoms", [vp, vp, vp]),
        ("pa_cif_strings", [vp, C.c_int32, C.c_char_p, C.c_size_t, szp]),
        ("pa_cif_types", [vp, vp, vp]),
        ("pa_read_forcefield", [C.c_char_p, C.POINTER(vp)]),
        ("pa_ff_info", [vp, i32p, dp, i32p]),
        ("pa_ff_params", [vp, vp, vp, vp, vp]),
        ("pa_ff_names", [vp, C.c_char_p, C.c_size_t, szp]),
        ("pa_ff_typeindex", [vp, C.c_char_p, i32p]),
        ("pa_read_guest", [C.c_char_p, vp, C.POINTER(vp)]),
        ("pa_guest_info", [vp, i32p, dp, dp, dp]),
        ("pa_guest_sites", [vp, vp, vp, vp]),
    ):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = args
    for name in ("pa_cif_destroy", "pa_ff_destroy", "pa_guest_destroy"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [vp]
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().pa_last_error(buf, len(buf))
    return buf.value.decode("utf-8", "replace")


def check(status: int) -> None:
    """Raises the exception class matching a C-ABI status."""
    if status == PA_OK:
        return
    msg = last_error()
    if status == PA_ERR_ARG:
        raise ArgumentError(msg)
    if status == PA_ERR_DIM:
        raise DimensionMismatch(msg)
    if status == PA_ERR_DEVICE:
        raise DeviceError(msg)
    raise RuntimeError(f"status {status}: {msg}")


def synchronize() -> None:
    """Waits for all device work (KernelAbstractions.synchronize(backend) in the reference)."""
    check(lib().pa_synchronize())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().pa_device_name(buf, len(buf)))
    return buf.value.decode()
