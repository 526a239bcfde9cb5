"""ctypes binding of libnimblesm_amd.so (the C ABI declared in
include/nimblesm_amd.h).  Loading fails loudly when the library has not been
built: there is no fallback implementation."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NSM_LIBRARY selects another build of the library (tuning builds); the
# in-tree build is the default
LIB_PATH = os.environ.get("NSM_LIBRARY") or os.path.join(HERE, "libnimblesm_amd.so")

NSM_OK, NSM_ERR_ARG, NSM_ERR_HIP, NSM_ERR_NEG_JACOBIAN, NSM_ERR_STATE, NSM_ERR_NOMEM = range(6)
NSM_FP64, NSM_MIXED = 0, 1
NSM_NOALLOC_MAX_SAMPLES = 4096  # include/nimblesm_amd.h
FIELDS = {
    "reference_coordinate": 0,
    "displacement": 1,
    "velocity": 2,
    "acceleration": 3,
    "internal_force": 4,
    "lumped_mass": 5,
    "deformation_gradient": 6,
    "stress": 7,
}
BC_KINDS = {"initial_velocity": 0, "prescribed_velocity": 1, "prescribed_displacement": 2}
KERNELS = {"element": 0, "node": 1, "reaction": 2, "closure": 3}

# every exported symbol with (restype, argtypes)
_P, _I64, _I, _D = C.c_void_p, C.c_int64, C.c_int, C.c_double
SIGNATURES = {
    "nsm_status_string": (C.c_char_p, [_I]),
    "nsm_last_error": (C.c_char_p, []),
    "nsm_version": (C.c_char_p, []),
    "nsm_hex_compute_deformation_gradients": (_I, [_I64, _P, _P, _P, _P]),
    "nsm_hex_compute_nodal_forces": (_I, [_I64, _P, _P, _P, _P, _P]),
    "nsm_hex_compute_lumped_mass": (_I, [_I64, _D, _P, _P, _P]),
    "nsm_hex_compute_characteristic_length": (_I, [_I64, _P, _P, _P]),
    "nsm_neohookean_get_stress": (_I, [_I64, _D, _D, _P, _P, _P]),
    "nsm_neohookean_get_stress_strided": (_I, [_I64, _I, _D, _D, _P, _I64, _I64, _I64, _P, _I64, _I64, _I64, _P]),
    "nsm_neohookean_get_stress_grouped": (_I, [_I64, _I64, _I, _D, _D, _P, _P, _P, _P, _P]),
    "nsm_model_create": (_I, [_I64, _P, _I, _P, _P, _P, _P, _P, _P, _I, C.POINTER(_P)]),
    "nsm_model_destroy": (_I, [_P]),
    "nsm_model_add_boundary_condition": (_I, [_P, _I, _I, _D, _I64, _P]),
    "nsm_model_apply_initial_conditions": (_I, [_P]),
    "nsm_model_apply_kinematic_conditions": (_I, [_P, _D, _D]),
    "nsm_model_compute_lumped_mass": (_I, [_P]),
    "nsm_model_get_critical_time_step": (_I, [_P, C.POINTER(_D)]),
    "nsm_model_compute_internal_force": (_I, [_P, _D, _D, _I]),
    "nsm_model_update_states": (_I, [_P]),
    "nsm_model_compute_deformation_gradients": (_I, [_P]),
    "nsm_model_compute_internal_force_from_stress": (_I, [_P]),
    "nsm_model_get_stream": (_I, [_P, C.POINTER(_P)]),
    "nsm_model_set_reaction_nodes": (_I, [_P, _I64, _P, _I]),
    "nsm_model_get_reaction_history": (_I, [_P, _I64, _I64, _P]),
    "nsm_model_sample_reaction_history": (_I, [_P, _I64, _I, _P]),
    "nsm_model_sample_reaction_history_noalloc": (_I, [_P, _I64, _I, _P]),
    "nsm_model_explicit_steps": (_I, [_P, _I64, _I64, _I64, _I64, _D, C.POINTER(_D)]),
    "nsm_model_get_field": (_I, [_P, _I, _P]),
    "nsm_model_set_field": (_I, [_P, _I, _P]),
    "nsm_model_field_device_pointer": (_I, [_P, _I, C.POINTER(_P), C.POINTER(_I)]),
    "nsm_model_set_uq_samples": (_I, [_P, _I, _P, _P, _I, _P]),
    "nsm_model_get_uq_field": (_I, [_P, _I, _I, _P]),
    "nsm_model_set_uq_field": (_I, [_P, _I, _I, _P]),
    "nsm_model_synchronize": (_I, [_P]),
    "nsm_model_enable_kernel_timing": (_I, [_P, _I]),
    "nsm_model_kernel_time": (_I, [_P, _I, C.POINTER(_D), C.POINTER(_I64)]),
    "nsm_model_get_element_order": (_I, [_P, _P]),
    "nsm_model_get_node_order": (_I, [_P, _P]),
    "nsm_model_assembly_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64),
                                       C.POINTER(_I64)]),
    "nsm_model_tile_elems": (_I, [_P, C.POINTER(_I)]),
    "nsm_model_info": (_I, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64)]),
    "nsm_ensemble_create": (_I, [_I, _I, C.POINTER(_P)]),
    "nsm_ensemble_destroy": (_I, [_P]),
    "nsm_ensemble_add_mesh": (_I, [_P, _I64, _P, _I, _P, _P, C.POINTER(_I)]),
    "nsm_ensemble_add_boundary_condition": (_I, [_P, _I, _I, _I, _I64, _P]),
    "nsm_ensemble_set_reaction_nodes": (_I, [_P, _I, _I64, _P, _I]),
    "nsm_ensemble_add_samples": (_I, [_P, _I, _I, _P, _P, _P, _P, C.POINTER(_I)]),
    "nsm_ensemble_set_samples": (_I, [_P, _I, _I, _P, _P, _P, _P]),
    "nsm_ensemble_setup": (_I, [_P]),
    "nsm_ensemble_get_critical_time_steps": (_I, [_P, _P]),
    "nsm_ensemble_run": (_I, [_P, _I64, _I64, _P, _I]),
    "nsm_ensemble_sample_reaction_history": (_I, [_P, _I, _P]),
    "nsm_ensemble_get_reaction_history": (_I, [_P, _I, _I64, _I64, _P]),
    "nsm_ensemble_get_field": (_I, [_P, _I, _I, _P]),
    "nsm_ensemble_info": (_I, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64),
                               C.POINTER(_I), C.POINTER(_I64)]),
    "nsm_ensemble_autotune_timings": (_I, [_P, _I, C.POINTER(_I), _P, _P]),
    "nsm_ensemble_enable_kernel_timing": (_I, [_P, _I]),
    "nsm_ensemble_kernel_time": (_I, [_P, _I, C.POINTER(_D), C.POINTER(_I64)]),
    "nsm_mlp_create": (_I, [_I, _P, C.c_uint64, C.POINTER(_P)]),
    "nsm_mlp_destroy": (_I, [_P]),
    "nsm_mlp_num_parameters": (_I, [_P, C.POINTER(_I64)]),
    "nsm_mlp_get_parameters": (_I, [_P, _P]),
    "nsm_mlp_set_parameters": (_I, [_P, _P]),
    "nsm_mlp_forward": (_I, [_P, _I64, _P, _P]),
    "nsm_mlp_loss_gradients": (_I, [_P, _I64, _P, _P, C.POINTER(_D), _P]),
    "nsm_mlp_train": (_I, [_P, _I64, _P, _P, _I, _D, _D, _P]),
    "nsm_jittered_cube": (_I, [_I, C.c_uint64, _D, _P, _P]),
}

_lib = None


class NimbleError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def lib():
    """Load the library (built by ``python -m nimblesm_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m nimblesm_amd.build)"
            )
        # A process must hold one HIP runtime.  PyTorch-ROCm ships its own
        # libamdhip64 (soname libamdhip64.so.7) and links it by an unversioned
        # name, so it has to be loaded first: our DT_NEEDED libamdhip64.so.7 then
        # resolves to that copy instead of pulling in a second runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != NSM_OK:
        raise NimbleError(status, lib().nsm_last_error().decode())
