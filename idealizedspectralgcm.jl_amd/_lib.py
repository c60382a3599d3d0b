"""ctypes binding of libjgcm_hip.so (declarations in include/jgcm.h).

The library is the only compute path: if it is missing or a GPU is absent the
calls raise; nothing here falls back to CPU arithmetic.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjgcm_hip.so")

JGCM_OK = 0
JGCM_ERR_ARG = -1
JGCM_ERR_HIP = -2
JGCM_ERR_STATE = -3
JGCM_ERR_NOMEM = -4
JGCM_ERR_UNSUPPORTED = -5

PREC = {"fp64": 0, "fp32": 1, "split32": 2, "bf16x3": 3}

TABLES = {
    "sinθ": 0, "cosθ": 1, "wts": 2, "λc": 3, "θc": 4, "λe": 5, "θe": 6,
    "qnm": 7, "dqnm": 8, "qwg": 9, "epsilon": 10, "laplacian_eig": 11,
    "coef_uvm": 12, "coef_uvc": 13, "coef_uvp": 14, "coef_dλ": 15,
    "coef_dθm": 16, "coef_dθp": 17, "coef_alp_a": 18, "coef_alp_b": 19,
}

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
SZ = ctypes.c_size_t

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "jgcm_last_error": [],
    "jgcm_device_count": [ctypes.POINTER(I)],
    "jgcm_malloc": [ctypes.POINTER(P), SZ],
    "jgcm_free": [P],
    "jgcm_memcpy_h2d": [P, P, SZ],
    "jgcm_memcpy_d2h": [P, P, SZ],
    "jgcm_copy": [P, P, P, SZ],
    "jgcm_copy_batch": [P, I, P, P, P],
    "jgcm_add_batch": [P, I, P, P, P],
    "jgcm_fill": [P, P, D, SZ],
    "jgcm_time_advance": [P, I, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                          ctypes.POINTER(SZ)],
    "jgcm_mesh_create": [I, I, I, I, I, D, ctypes.POINTER(P)],
    "jgcm_mesh_destroy": [P],
    "jgcm_mesh_lifetime": [P, ctypes.POINTER(I), ctypes.POINTER(I)],
    "jgcm_mesh_synchronize": [P],
    "jgcm_mesh_stream": [P],
    "jgcm_mesh_share_stream": [P, P],
    "jgcm_mesh_wait": [P, P],
    "jgcm_mesh_branch": [P, I],
    "jgcm_mesh_join": [P],
    "jgcm_mesh_set_precision": [P, I],
    "jgcm_mesh_get_precision": [P, ctypes.POINTER(I)],
    "jgcm_mesh_set_fft_shape": [P, I, I],
    "jgcm_mesh_get_fft_shape": [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)],
    "jgcm_mesh_set_legendre_shape": [P, I, I],
    "jgcm_mesh_get_legendre_shape": [P, ctypes.POINTER(I), ctypes.POINTER(I)],
    "jgcm_mesh_table_size": [P, I, ctypes.POINTER(SZ)],
    "jgcm_mesh_get_table": [P, I, P, SZ],
    "jgcm_trans_grid_to_spherical": [P, P, P, I],
    "jgcm_trans_grid_to_spherical_noalloc": [P, P, P, I],
    "jgcm_trans_spherical_to_grid": [P, P, P, I],
    "jgcm_trans_spherical_to_grid_batch": [P, I, ctypes.POINTER(P), ctypes.POINTER(P),
                                           ctypes.POINTER(P), ctypes.POINTER(I),
                                           ctypes.POINTER(I)],
    "jgcm_trans_grid_to_spherical_batch": [P, I, ctypes.POINTER(P), ctypes.POINTER(P),
                                           ctypes.POINTER(I), ctypes.POINTER(I)],
    "jgcm_trans_grid_to_fourier": [P, P, P, I],
    "jgcm_plan_spherical_to_grid_create": [I, P, P, P, P, P, P, P, I, ctypes.POINTER(P)],
    "jgcm_plan_grid_to_spherical_create": [I, P, P, P, P, P, P, I, ctypes.POINTER(P)],
    "jgcm_plan_execute": [P],
    "jgcm_plan_info": [P, ctypes.POINTER(I), ctypes.POINTER(I)],
    "jgcm_plan_work_list": [P, P, P],
    "jgcm_plan_destroy": [P],
    "jgcm_lpt_schedule": [I, P, I, P, P, ctypes.POINTER(I)],
    "jgcm_trans_grid_to_spherical_adjoint": [P, P, P, I],
    "jgcm_trans_spherical_to_grid_adjoint": [P, P, P, I],
    "jgcm_trans_spherical_to_grid_adjoint_batch": [P, I, ctypes.POINTER(P), ctypes.POINTER(P),
                                                   ctypes.POINTER(I), ctypes.POINTER(I), I],
    "jgcm_trans_grid_to_spherical_adjoint_batch": [P, I, ctypes.POINTER(P), ctypes.POINTER(P),
                                                   ctypes.POINTER(I)],
    "jgcm_vor_div_from_grid_uv_adjoint": [P, P, P, P, P, I],
    "jgcm_ucos_vcos_adjoint": [P, P, P, P, P, I],
    "jgcm_uv_grid_from_vor_div_adjoint": [P, P, P, P, P, I],
    "jgcm_vorticity_flux_adjoint": [P, P, P, P, D, P, P, P, P, P, I],
    "jgcm_damped_leapfrog_adjoint": [P, I, I, P, P, P, P, P, I],
    "jgcm_damped_leapfrog_adjoint_batch": [P, I, I, P, P, P, P, P, P, P],
    "jgcm_barotropic_step_adjoint": [P, I, P, P, P, D, P, P, I],
    "jgcm_sw_step_adjoint": [P, I, P, P, P, P, P, P, D, D, D, D, P, P, P, P, P, P, I],
    "jgcm_hs_column_adjoint": [P, P, D, P, P, P, P, P],
    "jgcm_hs_corrections_adjoint": [P, P, P, P, P, P, P, P, P, P, P, P, P],
    "jgcm_implicit_adjoint": [P, I, P, P, P, P, P],
    "jgcm_gradient_cos_adjoint": [P, P, P, P, I, I],
    "jgcm_divide_by_cos": [P, P, I],
    "jgcm_multiply_by_cos": [P, P, P, I],
    "jgcm_vor_div_from_grid_uv": [P, P, P, P, P, I],
    "jgcm_compute_alpha_operator": [P, P, P, D, P, I],
    "jgcm_uv_grid_from_vor_div": [P, P, P, P, P, I],
    "jgcm_compute_ucos_vcos_from_vor_div": [P, P, P, P, P, I],
    "jgcm_apply_laplacian": [P, P, I],
    "jgcm_compute_gradient_cos": [P, P, P, P, I],
    "jgcm_compute_gradients": [P, P, P, P, I],
    "jgcm_add_horizontal_advection": [P, P, P, P, P, I],
    "jgcm_area_weighted_global_mean": [P, P, ctypes.POINTER(D)],
    "jgcm_leapfrog_create": [P, D, I, D, D, ctypes.c_int64, I, ctypes.c_int64, ctypes.c_int64,
                             ctypes.POINTER(P)],
    "jgcm_leapfrog_destroy": [P],
    "jgcm_leapfrog_update_init_step": [P],
    "jgcm_leapfrog_state": [P, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int64)],
    "jgcm_leapfrog_set_time": [P, ctypes.c_int64],
    "jgcm_leapfrog_damping": [P, P, SZ],
    "jgcm_compute_spectral_damping": [P, P, P, P, I],
    "jgcm_filtered_leapfrog": [P, P, P, P, P, I],
    "jgcm_damped_leapfrog": [P, P, P, P, P, I],
    "jgcm_damped_leapfrog_batch": [P, I] + [ctypes.POINTER(P)] * 4 + [ctypes.POINTER(I)] * 2,
    "jgcm_sw_implicit_correction": [P, P, P, P, P, D, P, P, I],
    "jgcm_vorticity_flux": [P, P, P, P, D, P, P, P, I, I],
    "jgcm_axpby": [P, SZ, D, P, D, P, P],
    "jgcm_xmy": [P, SZ, D, P, P, P, I],
    "jgcm_kinetic_energy": [P, P, P, P, P, I],
    "jgcm_shallow_water_physics": [P, P, P, P, P, D, D, P, P, P, I],
    "jgcm_sw_explicit": [P, P, P, P, P, D, D, P, D, P, P, P, P, P, P, P, P, I],
    "jgcm_sw_height_tendency": [P, P, P, P, P, P, P, I],
    "jgcm_sw_height_tendency_carried": [P, P, P, P, P, P, P, P, I],
    "jgcm_sw_spectral_step": [P, D, P, P, P, P, P, ctypes.POINTER(P), ctypes.POINTER(P),
                              ctypes.POINTER(P), I],
    "jgcm_sw_spectral_step_ex": [P, D, P, P, P, P, P, ctypes.POINTER(P), ctypes.POINTER(P),
                                 ctypes.POINTER(P), I, I],
    "jgcm_map": [P, I, SZ, P, D, D, P],
    "jgcm_vert_coord_create": [P, I, I, I, I, D, I, ctypes.POINTER(P)],
    "jgcm_vert_coord_create_ex": [P, I, I, I, I, D, I, D, D, D, D, D, ctypes.POINTER(P)],
    "jgcm_vert_coord_destroy": [P],
    "jgcm_vert_coord_get": [P, P, P],
    "jgcm_pressure_variables": [P, P, P, P, P, P, P],
    "jgcm_compute_pressures_and_heights": [P, P, P, P, P, P, P, P, P, P, P, P],
    "jgcm_compute_geopotential": [P, P, P, P, P, P, P, P],
    "jgcm_vert_advection": [P, P, P, P, P],
    "jgcm_four_in_one": [P, P] + [P] * 17,
    "jgcm_semi_implicit_create": [P, P, P, D, P, ctypes.POINTER(P)],
    "jgcm_semi_implicit_destroy": [P],
    "jgcm_semi_implicit_update_init_step": [P],
    "jgcm_semi_implicit_get": [P, I, P, SZ],
    "jgcm_implicit_correction": [P] * 10,
    "jgcm_hs_energy_implicit_correction": [P] * 11,
    "jgcm_hs_energy_implicit_leapfrog": [P] * 18,
    "jgcm_hs_physics": [P, P, ctypes.c_int64] + [P] * 15 + [I],
    "jgcm_hs_column_dynamics": [P] * 27,
    "jgcm_hs_physics_column_dynamics": [P, P, ctypes.c_int64] + [P] * 27,
    "jgcm_hs_physics_column_dynamics_ex": [P, P, ctypes.c_int64] + [P] * 29 + [I],
    "jgcm_hs_corrections": [P] * 8,
    "jgcm_hs_corrections_advance": [P, P, P, P, P, P],
    "jgcm_hs_implicit_leapfrog_advance": [P] * 6 + [ctypes.POINTER(P)] * 4,
    "jgcm_hs_scalars": [P, P, I],
    "jgcm_hs_scalars_device": [P, P, I],
    "jgcm_zonal_mean_accumulate": [P, P, P, P, P, P, I, P, P, P, P, P],
    "jgcm_mesh_reserve": [P, I],
    "jgcm_graph_begin": [P],
    "jgcm_graph_end": [P, ctypes.POINTER(P)],
    "jgcm_graph_launch": [P],
    "jgcm_graph_destroy": [P],
}
_RESTYPE = {"jgcm_last_error": ctypes.c_char_p, "jgcm_mesh_stream": P}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() "
                "(the HIP library is the only compute path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, I)
        _lib = L
    return _lib


class JGCMError(RuntimeError):
    pass


def check(rc):
    """Map a status code to the reference's exception kinds: argument/shape
    failures -> AssertionError (reference @assert), the rest -> JGCMError."""
    if rc == JGCM_OK:
        return
    msg = lib().jgcm_last_error().decode(errors="replace")
    if rc == JGCM_ERR_ARG:
        raise AssertionError(msg)
    raise JGCMError(f"jgcm error {rc}: {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args))
