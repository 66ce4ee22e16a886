"""ctypes binding of ``libspecfem3d_amd.so`` -- the same C-ABI the Fortran host binds.

Every entry point takes its arguments by reference (Fortran calling convention), so this
module is the Python counterpart of the ``call ...(Mesh_pointer, ...)`` statements in
``src/specfem3D/prepare_gpu.f90`` and the ``*_calling`` routines. Loading fails loudly if the
library has not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspecfem3d_amd.so")
_lib = None

# entry points declared in include/specfem3d_amd.h (checked by the CPU test-suite)
EXPORTED = [
    "initialize_gpu_device_", "prepare_constants_device_", "prepare_fields_elastic_device_",
    "prepare_fields_acoustic_device_", "prepare_fields_elastic_adj_dev_", "prepare_fields_acoustic_adj_dev_",
    "prepare_sim2_or_3_const_device_", "prepare_cleanup_device_",
    "update_displacement_cuda_", "update_displacement_ac_cuda_", "kernel_3_a_cuda_", "kernel_3_b_cuda_",
    "kernel_3_acoustic_cuda_", "compute_forces_viscoelastic_cuda_", "compute_forces_acoustic_cuda_",
    "acoustic_enforce_free_surf_cuda_", "compute_stacey_viscoelastic_cuda_",
    "compute_stacey_viscoelastic_undoatt_cuda_", "compute_stacey_acoustic_cuda_",
    "compute_stacey_acoustic_undoatt_cuda_", "compute_coupling_ac_el_cuda_", "compute_coupling_el_ac_cuda_",
    "compute_add_sources_el_cuda_", "compute_add_sources_el_s3_cuda_", "add_sources_el_sim_type_2_or_3_",
    "compute_add_sources_ac_cuda_", "compute_add_sources_ac_s3_cuda_", "add_sources_ac_sim_2_or_3_cuda_",
    "compute_seismograms_cuda_", "compute_kernels_elastic_cuda_", "compute_strain_cuda_",
    "compute_kernels_acoustic_cuda_", "get_norm_elastic_from_device_", "get_norm_acoustic_from_device_",
    "transfer_fields_el_to_device_", "transfer_fields_el_from_device_", "transfer_b_fields_to_device_",
    "transfer_b_fields_from_device_", "transfer_displ_to_device_", "transfer_displ_from_device_",
    "transfer_veloc_to_device_", "transfer_veloc_from_device_", "transfer_accel_to_device_",
    "transfer_accel_from_device_", "transfer_b_displ_to_device_", "transfer_b_displ_from_device_",
    "transfer_b_veloc_to_device_", "transfer_b_veloc_from_device_", "transfer_b_accel_to_device_",
    "transfer_b_accel_from_device_", "transfer_fields_ac_to_device_", "transfer_fields_ac_from_device_",
    "transfer_b_fields_ac_to_device_", "transfer_b_fields_ac_from_device_", "transfer_potential_ac_from_device_",
    "transfer_b_potential_ac_from_device_", "transfer_b_potential_ac_to_device_",
    "transfer_b_potential_dot_dot_ac_from_device_", "transfer_b_potential_dot_dot_ac_to_device_",
    "transfer_dot_dot_to_device_", "transfer_dot_dot_from_device_", "transfer_b_dot_dot_from_device_",
    "transfer_kernels_el_to_host_", "transfer_kernels_ac_to_host_", "transfer_strain_from_device_",
    "transfer_b_strain_to_device_", "transfer_rmemory_from_device_", "transfer_b_rmemory_to_device_",
    "transfer_boun_pot_from_device_", "transfer_asmbl_pot_to_device_", "transfer_boun_accel_from_device_",
    "transfer_boundary_from_device_a_", "transfer_boundary_to_device_a_", "transfer_asmbl_accel_to_device_",
    "transfer_sync_accel_to_device_", "sync_copy_from_device_", "get_free_device_memory_",
    "output_free_device_memory_", "register_host_array_", "unregister_host_array_",
    "prepare_fields_elastic_pml_", "pml_impose_boundary_condition_elastic_cuda_", "transfer_pml_displ_from_device_",
    "transfer_pml_displ_to_device_",
    "sf_set_options", "sf_set_tuning", "sf_set_element_order", "sf_color_elements", "sf_batch_point_flags", "sf_snapshot_save", "sf_snapshot_load",
    "sf_history_reserve", "sf_history_store", "sf_history_load", "sf_checkpoints_free", "sf_zero_kernels",
    "sf_batch_create", "sf_batch_add_source", "sf_batch_add_receiver", "sf_batch_finalize", "sf_batch_set_options",
    "sf_batch_get_options", "sf_batch_set_tuning", "sf_batch_get_tuning", "sf_batch_autotune", "sf_batch_run", "sf_batch_current_step", "sf_batch_seismograms",
    "sf_batch_displ", "sf_batch_potential", "sf_batch_reset", "sf_batch_time", "sf_batch_destroy", "sf_set_jacobian", "sf_synchronize", "sf_last_error", "sf_norm_noalloc",
]


def load(path: str | None = None):
    """Load the HIP library; raises if it is missing (build it with ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} not found: the HIP library is not built (run __graft_entry__.build())")
    _lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
    _lib.sf_last_error.restype = C.c_char_p
    return _lib


class Args:
    """Keeps ctypes objects alive for by-reference calls."""

    def __init__(self):
        self.keep = []

    def i(self, v) -> C.POINTER(C.c_int):
        o = C.c_int(int(v))
        self.keep.append(o)
        return C.byref(o)

    def f(self, v):
        o = C.c_float(float(v))
        self.keep.append(o)
        return C.byref(o)

    def arr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.size == 0:
            a = np.zeros(1, dtype=dtype)
        self.keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    def fa(self, a):
        return self.arr(a, np.float32)

    def ia(self, a):
        return self.arr(a, np.int32)

    def da(self, a):
        return self.arr(a, np.float64)


def ptr(a: np.ndarray):
    """Pointer to an existing (contiguous) numpy array -- caller keeps it alive."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
