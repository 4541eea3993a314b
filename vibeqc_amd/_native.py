"""ctypes binding of ``libvibeqc_amd.so`` (the HIP engine's C ABI).

The structures mirror ``include/generativeqc/{generativeqc,fock}.h`` field for
field (the reference binds the same ABI through ctypes in
python/generativeqc/_native.py:1009-1030 and python/generativeqc/fock.py:104-224).
There is no fallback: if the library is missing, loading raises.
"""

from __future__ import annotations

import ctypes as ct
import os
import threading
from pathlib import Path

ABI_VERSION = 0

STATUS = {
    0: "SUCCESS",
    1: "INVALID_ARGUMENT",
    2: "ABI_MISMATCH",
    3: "NOT_IMPLEMENTED",
    4: "NOT_CONVERGED",
    5: "NUMERICAL_FAILURE",
    6: "DEVICE_ERROR",
    7: "OUT_OF_MEMORY",
    8: "INTERNAL_ERROR",
    9: "PRECISION_UNAVAILABLE",
}

BACKEND_CPU_REFERENCE = 0
BACKEND_HIP = 1
BASIS_CARTESIAN = 0
BASIS_SPHERICAL = 1
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2}

_LIBRARY_PATH = Path(os.environ.get("VIBEQC_AMD_LIBRARY", Path(__file__).resolve().parent / "libvibeqc_amd.so"))
_lock = threading.Lock()
_library = None


class NativeError(RuntimeError):
    """A non-success status from the native library."""

    def __init__(self, status: int, detail: str):
        self.status = status
        self.status_name = STATUS.get(status, str(status))
        super().__init__(f"{self.status_name}: {detail}")


class NotConvergedError(NativeError):
    pass


class ContextDescriptor(ct.Structure):
    _fields_ = [("struct_size", ct.c_uint32), ("abi_version", ct.c_uint32), ("device_id", ct.c_int32), ("backend", ct.c_int32)]


class Atom(ct.Structure):
    _fields_ = [("atomic_number", ct.c_int32), ("x", ct.c_double), ("y", ct.c_double), ("z", ct.c_double)]


class Primitive(ct.Structure):
    _fields_ = [("exponent", ct.c_double), ("coefficient", ct.c_double)]


class Shell(ct.Structure):
    _fields_ = [
        ("atom_index", ct.c_uint32),
        ("angular_momentum", ct.c_uint32),
        ("primitive_offset", ct.c_uint32),
        ("primitive_count", ct.c_uint32),
    ]


class SystemDescriptor(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("atoms", ct.POINTER(Atom)),
        ("atom_count", ct.c_uint32),
        ("shells", ct.POINTER(Shell)),
        ("shell_count", ct.c_uint32),
        ("primitives", ct.POINTER(Primitive)),
        ("primitive_count", ct.c_uint32),
        ("charge", ct.c_int32),
        ("multiplicity", ct.c_uint32),
        ("basis_representation", ct.c_int32),
    ]


class FockTermC(ct.Structure):
    _fields_ = [
        ("present", ct.c_int32),
        ("coefficient", ct.c_double),
        ("op", ct.c_int32),
        ("omega", ct.c_double),
        ("approximation", ct.c_int32),
    ]


class FockSpecC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("spec_version", ct.c_uint32),
        ("spin", ct.c_int32),
        ("derivative_order", ct.c_uint32),
        ("coulomb", FockTermC),
        ("exchange", FockTermC),
    ]


class FockControlsC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("screening_tolerance", ct.c_double),
        ("metric_relative_threshold", ct.c_double),
        ("device_budget_bytes", ct.c_uint64),
    ]


_DP = ct.POINTER(ct.c_double)


class FockResultC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("matrix_count", ct.c_uint64),
        ("gradient_count", ct.c_uint64),
        ("coulomb", _DP),
        ("exchange_alpha", _DP),
        ("exchange_beta", _DP),
        ("fock_alpha", _DP),
        ("fock_beta", _DP),
        ("gradient", _DP),
        ("energy_one_electron", ct.c_double),
        ("energy_two_electron", ct.c_double),
        ("nuclear_repulsion", ct.c_double),
    ]


class ScfControlsC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("max_iterations", ct.c_uint32),
        ("diis_history", ct.c_uint32),
        ("energy_tolerance", ct.c_double),
        ("density_tolerance", ct.c_double),
    ]


class ScfResultC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("density", _DP),
        ("forces", _DP),
        ("density_count", ct.c_uint64),
        ("force_count", ct.c_uint64),
        ("energy", ct.c_double),
        ("energy_change", ct.c_double),
        ("density_rms", ct.c_double),
        ("iterations", ct.c_uint32),
        ("initial_density_used", ct.c_int32),
        ("fock_builds", ct.c_uint64),
    ]


class FockDiagnosticC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("requested", FockSpecC),
        ("resolved", FockSpecC),
        ("backend", ct.c_int32),
        ("resolved_schedule", ct.c_int32),
        ("source_schedule", ct.c_int32),
        ("nbf", ct.c_uint64),
        ("coordinate_count", ct.c_uint64),
        ("device_bytes", ct.c_uint64),
        ("device_budget_bytes", ct.c_uint64),
        ("auxiliary_rank", ct.c_uint64),
        ("auxiliary_tile", ct.c_uint64),
        ("screening_tolerance", ct.c_double),
        ("metric_relative_threshold", ct.c_double),
        ("df_streamed", ct.c_int32),
        ("direct_schedule", ct.c_char * 96),
        ("df_value_backend", ct.c_char * 48),
        ("df_value_mapping", ct.c_char * 32),
        ("df_response_mapping", ct.c_char * 32),
        ("one_electron_value_backend", ct.c_char * 32),
        ("one_electron_value_mapping", ct.c_char * 32),
        ("one_electron_response_mapping", ct.c_char * 32),
    ]


METHOD_RHF = 1
METHOD_UHF = 2
DENSITY_FITTING_NONE = 0
DENSITY_FITTING_CUDA = 2
PRECISION_MODE = {"fp64": 0, "auto": 1}
BATCH_WARM_STARTS = 1


class MethodDescriptorC(ct.Structure):
    """generativeqc_method_descriptor (reference generativeqc.h:723-769)."""

    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("method", ct.c_int32),
        ("max_iterations", ct.c_uint32),
        ("diis_history", ct.c_uint32),
        ("energy_tolerance", ct.c_double),
        ("density_tolerance", ct.c_double),
        ("screening_tolerance", ct.c_double),
        ("density_fitting_mode", ct.c_int32),
        ("density_fitting_auxiliary_basis", ct.c_void_p),
        ("density_fitting_relative_threshold", ct.c_double),
        ("density_fitting_memory_budget_bytes", ct.c_uint64),
        ("precision_mode", ct.c_int32),
        ("correlation_memory_budget_bytes", ct.c_uint64),
        ("mp2_denominator_threshold", ct.c_double),
        ("ks_options", ct.c_void_p),
        ("ccsd_max_iterations", ct.c_uint32),
        ("ccsd_diis_history", ct.c_uint32),
        ("ccsd_energy_tolerance", ct.c_double),
        ("ccsd_residual_tolerance", ct.c_double),
        ("ccsd_denominator_threshold", ct.c_double),
        ("ccsd_damping", ct.c_double),
        ("ccsd_level_shift", ct.c_double),
        ("ccsd_frozen_core", ct.c_uint32),
        ("initial_guess", ct.c_void_p),
    ]


class BatchInputC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("coordinates", _DP),
        ("coordinate_count", ct.c_uint32),
    ]


class BatchItemResultC(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("status", ct.c_int32),
        ("energy", ct.c_double),
        ("forces", _DP),
        ("force_count", ct.c_uint32),
        ("iterations", ct.c_uint32),
        ("energy_change", ct.c_double),
        ("density_rms", ct.c_double),
        ("converged", ct.c_int32),
        ("executed_backend", ct.c_int32),
        ("bucket_id", ct.c_uint32),
        ("warm_start_used", ct.c_int32),
        ("warm_start_fallback", ct.c_int32),
    ]


class HfWarmState(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("density", _DP),
        ("density_count", ct.c_uint64),
        ("coordinates", _DP),
        ("coordinate_count", ct.c_uint64),
        ("energy", ct.c_double),
        ("energy_change", ct.c_double),
        ("density_rms", ct.c_double),
        ("iterations", ct.c_int32),
        ("present", ct.c_int32),
    ]


class ResultDescriptorC(ct.Structure):
    """generativeqc_result_descriptor (generativeqc.h:1067-1081)."""

    _fields_ = [
        ("struct_size", ct.c_uint32),
        ("abi_version", ct.c_uint32),
        ("energy", ct.c_double),
        ("forces", _DP),
        ("force_count", ct.c_uint32),
        ("iterations", ct.c_uint32),
        ("energy_change", ct.c_double),
        ("density_rms", ct.c_double),
        ("converged", ct.c_int32),
        ("executed_backend", ct.c_int32),
    ]


def descriptor(kind, **values):
    """A versioned descriptor with struct_size / abi_version filled in."""
    return kind(struct_size=ct.sizeof(kind), abi_version=ABI_VERSION, **values)


EXPORTS = {
    "generativeqc_get_abi_version": (ct.c_uint32, []),
    "generativeqc_get_source_identity": (ct.c_char_p, []),
    "generativeqc_status_message": (ct.c_char_p, [ct.c_int32]),
    "generativeqc_context_create": (ct.c_int32, [ct.POINTER(ContextDescriptor), ct.POINTER(ct.c_void_p)]),
    "generativeqc_context_destroy": (None, [ct.c_void_p]),
    "generativeqc_context_get_last_detail": (ct.c_char_p, [ct.c_void_p]),
    "generativeqc_context_last_error": (ct.c_char_p, [ct.c_void_p]),
    "generativeqc_system_create": (ct.c_int32, [ct.c_void_p, ct.POINTER(SystemDescriptor), ct.POINTER(ct.c_void_p)]),
    "generativeqc_system_destroy": (None, [ct.c_void_p]),
    "generativeqc_fock_plan_create": (
        ct.c_int32,
        [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(FockSpecC), ct.POINTER(FockControlsC), ct.POINTER(ct.c_void_p)],
    ),
    "generativeqc_fock_plan_destroy": (None, [ct.c_void_p]),
    "generativeqc_fock_plan_last_error": (ct.c_char_p, [ct.c_void_p]),
    "generativeqc_fock_plan_evaluate": (
        ct.c_int32,
        [ct.c_void_p, _DP, ct.c_uint64, _DP, ct.c_uint64, ct.POINTER(FockResultC)],
    ),
    "generativeqc_fock_plan_solve": (
        ct.c_int32,
        [ct.c_void_p, ct.POINTER(ScfControlsC), _DP, ct.c_uint64, ct.POINTER(ScfResultC)],
    ),
    "generativeqc_fock_plan_diagnostic": (ct.c_int32, [ct.c_void_p, ct.POINTER(FockDiagnosticC)]),
    "vibeqc_fock_plan_set_precision": (ct.c_int32, [ct.c_void_p, ct.c_int32]),
    "vibeqc_fock_plan_evaluate_occupied": (ct.c_int32, [ct.c_void_p, _DP, ct.c_uint64, ct.c_uint64, _DP, _DP]),
    "vibeqc_fock_plan_evaluate_occupied_noalloc": (
        ct.c_int32,
        [ct.c_void_p, ct.c_void_p, ct.c_uint64, ct.c_uint64, ct.c_void_p, ct.c_void_p],
    ),
    "vibeqc_fock_plan_benchmark_occupied": (
        ct.c_int32,
        [ct.c_void_p, _DP, ct.c_uint64, ct.c_uint64, ct.c_uint32, _DP],
    ),
    "vibeqc_fock_plan_shape": (ct.c_int32, [ct.c_void_p, ct.POINTER(ct.c_uint64)]),
    "vibeqc_fock_plan_streaming": (ct.c_int32, [ct.c_void_p, ct.POINTER(ct.c_uint64)]),
    "vibeqc_fock_plan_last_solve_stats": (ct.c_int32, [ct.c_void_p, ct.POINTER(ct.c_uint64), ct.c_uint64]),
    "vibeqc_fock_plan_autotune": (
        ct.c_int32,
        [ct.c_void_p, _DP, ct.c_uint64, ct.c_uint64, ct.c_uint32, ct.POINTER(ct.c_int32), _DP],
    ),
    "vibeqc_fock_plan_launch_shape": (ct.c_int32, [ct.c_void_p, ct.c_uint64, ct.POINTER(ct.c_int32)]),
    "vibeqc_tuning_record": (
        ct.c_int32,
        [ct.c_char_p, ct.c_int32, ct.c_uint64, ct.c_uint64, ct.c_uint64, ct.c_int32, ct.c_int32, ct.c_double],
    ),
    "vibeqc_tuning_lookup": (
        ct.c_int32,
        [ct.c_char_p, ct.c_int32, ct.c_uint64, ct.c_uint64, ct.c_uint64, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32)],
    ),
    "vibeqc_tuning_count": (ct.c_int32, [ct.POINTER(ct.c_uint64)]),
    "vibeqc_tuning_load": (ct.c_int32, [ct.c_char_p, ct.POINTER(ct.c_int32)]),
    "vibeqc_tuning_save": (ct.c_int32, [ct.c_char_p]),
    "vibeqc_tuning_clear": (ct.c_int32, []),
    "vibeqc_fock_plan_one_electron": (ct.c_int32, [ct.c_void_p, _DP, _DP, ct.c_uint64]),
    "generativeqc_batch_prepare": (
        ct.c_int32,
        [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.c_uint32, ct.POINTER(MethodDescriptorC), ct.c_uint32, ct.POINTER(ct.c_void_p)],
    ),
    "vibeqc_batch_prepare": (
        ct.c_int32,
        [
            ct.c_void_p,
            ct.POINTER(ct.c_void_p),
            ct.POINTER(ct.c_void_p),
            ct.c_uint32,
            ct.POINTER(MethodDescriptorC),
            ct.c_uint32,
            ct.POINTER(ct.c_void_p),
        ],
    ),
    "generativeqc_batch_destroy": (None, [ct.c_void_p]),
    "generativeqc_batch_get_system_count": (ct.c_uint32, [ct.c_void_p]),
    "generativeqc_batch_execute": (
        ct.c_int32,
        [ct.c_void_p, ct.POINTER(BatchInputC), ct.c_uint32, ct.POINTER(BatchItemResultC), ct.c_uint32],
    ),
    "generativeqc_batch_set_warm_start_updates": (ct.c_int32, [ct.c_void_p, ct.c_int32]),
    "generativeqc_batch_get_last_fock_builds": (ct.c_int32, [ct.c_void_p, ct.c_uint32, ct.POINTER(ct.c_uint64)]),
    "vibeqc_batch_bucket": (ct.c_int32, [ct.c_void_p, ct.c_uint32, ct.POINTER(ct.c_uint32)]),
    "generativeqc_batch_get_hf_warm_state": (ct.c_int32, [ct.c_void_p, ct.c_uint32, ct.POINTER(HfWarmState)]),
    "generativeqc_batch_restore_hf_warm_states": (ct.c_int32, [ct.c_void_p, ct.POINTER(HfWarmState), ct.c_uint32]),
    "generativeqc_batch_clear_warm_starts": (ct.c_int32, [ct.c_void_p]),
    "generativeqc_calculation_prepare": (
        ct.c_int32,
        [ct.c_void_p, ct.c_void_p, ct.POINTER(MethodDescriptorC), ct.POINTER(ct.c_void_p)],
    ),
    "generativeqc_calculation_destroy": (None, [ct.c_void_p]),
    "generativeqc_calculation_execute": (ct.c_int32, [ct.c_void_p, ct.POINTER(ResultDescriptorC)]),
    "vibeqc_calculation_set_precision": (ct.c_int32, [ct.c_void_p, ct.c_int32]),
    "vibeqc_calculation_last_state": (ct.c_int32, [ct.c_void_p, _DP, ct.c_uint64, ct.POINTER(ct.c_uint64)]),
}


def library_path() -> Path:
    return _LIBRARY_PATH


def load():
    """Load and bind the engine; raises if it has not been built."""
    global _library
    with _lock:
        if _library is None:
            if not _LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"HIP engine not built: {_LIBRARY_PATH} is missing "
                    "(run `python -m vibeqc_amd.build` or __graft_entry__.build())"
                )
            lib = ct.CDLL(str(_LIBRARY_PATH))
            for name, (restype, argtypes) in EXPORTS.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _library = lib
        return _library


def check(status: int, detail) -> None:
    if status == 0:
        return
    text = detail.decode() if isinstance(detail, bytes) else str(detail or "")
    if status == 4:
        raise NotConvergedError(status, text)
    raise NativeError(status, text)


def pointer(array):
    return None if array is None else array.ctypes.data_as(_DP)
