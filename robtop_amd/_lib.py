"""ctypes binding of librobtop_amd.so (the C ABI in include/robtop_amd.h).

The shared library is loaded from this package directory only. There is no
fallback: if it is missing or fails to load, every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librobtop_amd.so")
_LIB = None

RT_OK, RT_EARG, RT_EHIP, RT_ESTATE, RT_ENUM, RT_EINTERNAL = 0, -1, -2, -3, -4, -5
RT_SOLVE_FP64, RT_SOLVE_FP32, RT_SOLVE_MIXED, RT_SOLVE_FP32_IR, RT_SOLVE_RELIABLE = 0, 1, 2, 3, 4
RT_VEC_U, RT_VEC_F, RT_VEC_R = 0, 1, 2


class RobtopError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"robtop_amd error {code}: {msg}")
        self.code = code


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("rel_res", ctypes.c_double),
                ("rel_res_pcg", ctypes.c_double), ("seconds", ctypes.c_double),
                ("converged", ctypes.c_int)]


class LanczosStats(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_int), ("solves", ctypes.c_int), ("pcg_iterations", ctypes.c_int)]


P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
F = ctypes.c_float
SZ = ctypes.c_size_t

# name: (argtypes, restype)
SIGNATURES = {
    "rt_last_error": ([], ctypes.c_char_p),
    "rt_version": ([], ctypes.c_char_p),
    "rt_create_from_voxels": ([I, P, SZ, P], I),
    "rt_destroy": ([P], I),
    "rt_num_levels": ([P], I),
    "rt_num_layers": ([P], I),
    "rt_layer_active_elements": ([P, I], I),
    "rt_level_info": ([P, I, P], I),
    "rt_level_masks": ([P, I, P, P], I),
    "rt_num_active": ([P, I, P, P], I),
    "rt_surface_nodes_box": ([P, I, P], I),
    "rt_set_material": ([P, D, D, F], I),
    "rt_get_template_matrix": ([P, P], I),
    "rt_set_support_box": ([P, P], I),
    "rt_set_support_active": ([P, P], I),
    "rt_set_force_box": ([P, P], I),
    "rt_set_force_active": ([P, P], I),
    "rt_get_force_box": ([P, P], I),
    "rt_set_shell_box": ([P, P], I),
    "rt_init_density": ([P, F], I),
    "rt_set_density_box": ([P, P], I),
    "rt_get_density_box": ([P, P], I),
    "rt_set_density_active": ([P, P], I),
    "rt_get_density_active": ([P, P], I),
    "rt_get_rhop_box": ([P, P], I),
    "rt_update_stencil": ([P], I),
    "rt_get_stencil_box": ([P, I, P], I),
    "rt_set_vcycle_precision": ([P, I], I),
    "rt_set_level_vec": ([P, I, I, P], I),
    "rt_get_level_vec": ([P, I, I, P], I),
    "rt_gs_relax": ([P, I, I, I], I),
    "rt_update_residual": ([P, I], I),
    "rt_apply_k": ([P, P, P], I),
    "rt_restrict_residual": ([P, I], I),
    "rt_prolongate_correction": ([P, I], I),
    "rt_reset_displacement": ([P, I], I),
    "rt_solve_coarsest": ([P], I),
    "rt_relative_residual": ([P, P], I),
    "rt_vcycle": ([P, I, I, P], I),
    "rt_solve": ([P, D, I, I, I, I, P], I),
    "rt_set_displacement_box": ([P, P], I),
    "rt_get_displacement_box": ([P, P], I),
    "rt_get_displacement_active": ([P, P], I),
    "rt_compliance": ([P, P], I),
    "rt_compute_sensitivity": ([P], I),
    "rt_filter_sensitivity": ([P, D], I),
    "rt_get_sensitivity_box": ([P, P], I),
    "rt_set_sensitivity_box": ([P, P], I),
    "rt_update_densities": ([P, F, F, F, F, P, P], I),
    "rt_volume": ([P, P], I),
    "rt_voxelize_mesh": ([P, I, P, I, I, P, SZ, P, P], I),
    "rt_solve_batch": ([P, I, P, P, D, I, I, I, P, P], I),
    "rt_set_block_solves": ([P, I], I),
    "rt_set_coarse_direct_dofs": ([P, I], I),
    "rt_set_level_sweeps": ([P, I, I], I),
    "rt_preconditioner_levels": ([P, ctypes.POINTER(I)], I),
    "rt_voxelize_mesh_noalloc": ([P, I, P, I, I, P, SZ, P, P], I),
    "rt_set_preconditioner": ([P, I, I], I),
    "rt_set_graphs": ([P, I], I),
    "rt_mma_create": ([I, I, P, P, P], I),
    "rt_mma_destroy": ([P], I),
    "rt_mma_update": ([P, P, P, P, P], I),
    "rt_mma_update_rows": ([P, P, P, P, P], I),
    "rt_mma_get_x": ([P, P], I),
    "rt_mma_stats": ([P, P, P, P, P], I),
    "rt_set_load_nodes": ([P, P, P], I),
    "rt_num_load_nodes": ([P, P], I),
    "rt_force_project": ([P], I),
    "rt_rand_force": ([P, ctypes.c_uint64], I),
    "rt_modified_pm": ([P, I, P, P], I),
    "rt_set_coarse_stencil_bits": ([P, I], I),
    "rt_set_fine_smoother": ([P, I], I),
    "rt_set_warm_start": ([P, I], I),
    "rt_set_island_deflation": ([P, F, I], I),
    "rt_num_islands": ([P, P], I),
    "rt_block_lanczos": ([P, I, I, D, D, I, I, I, ctypes.c_uint64, P, P], I),
    "rt_net_create": ([ctypes.c_uint64, P], I),
    "rt_net_destroy": ([P], I),
    "rt_net_num_params": ([P, P], I),
    "rt_net_get_params": ([P, P], I),
    "rt_net_set_params": ([P, P], I),
    "rt_net_forward": ([P, P, P, P], I),
    "rt_net_gradient": ([P, P, P, P, P, P], I),
    "rt_net_add_sample": ([P, P], I),
    "rt_net_evaluate": ([P, P, P], I),
    "rt_net_predict_coarse": ([P, P], I),
    "rt_net_time_forward": ([P, P, I, P], I),
    "rt_net_num_samples": ([P, P], I),
    "rt_net_clear_samples": ([P], I),
    "rt_net_train": ([P, I, I, F, ctypes.c_uint64, P], I),
    "rt_net_adam_step": ([P, P, F], I),
    "rt_set_learned_warm_start": ([P, P], I),
    "rt_learned_warm_start_info": ([P, P], I),
    "rt_autotune": ([P, I, P], I),
    "rt_cg_ritz_values": ([P, I, P, P, P], I),
    "rt_cg_coefficients": ([P, I, P, P, P], I),
    "rt_synchronize": ([P], I),
    "rt_time_vcycle": ([P, I, I, I, I, P], I),
    "rt_time_vcycle_batch": ([P, I, I, I, I, P], I),
    "rt_time_fine_sweep": ([P, I, I, P], I),
    "rt_time_fine_residual": ([P, I, I, P], I),
}


def lib():
    """Load librobtop_amd.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m robtop_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(code):
    if code < 0:
        raise RobtopError(code, lib().rt_last_error().decode())
    return code


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def as_c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)
