"""ctypes mirror of include/pmg.h and of the problem generator's config struct.

Loading fails loudly: there is no Python fallback for any kernel.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_DIR = PKG / "lib"

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)

PMG_OK = 0
STATUS = {
    0: "PMG_OK", 1: "PMG_ERR_INVALID_ARGUMENT", 2: "PMG_ERR_DEVICE", 3: "PMG_ERR_OUT_OF_MEMORY",
    4: "PMG_ERR_NOT_READY", 5: "PMG_ERR_SINGULAR_PATCH", 6: "PMG_ERR_SINGULAR_COARSE",
    7: "PMG_ERR_NONFINITE", 8: "PMG_ERR_DIVERGED", 9: "PMG_ERR_NOT_CONVERGED", 10: "PMG_ERR_UNSUPPORTED",
}
PREC_FP64, PREC_FP32, PREC_FP32_BF16 = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1


class PmgParams(C.Structure):
    _fields_ = [
        ("sc_smoother_saddle", C.c_int),
        ("sc_pre_smooth_saddle", C.c_int),
        ("sc_post_smooth_saddle", C.c_int),
        ("sc_smooth_damp_factor_saddle", C.c_double),
        ("sc_smooth_damp_factor_coarse_saddle", C.c_double),
        ("sc_gmg_damp_factor_saddle", C.c_double),
        ("sc_coarse_smoother_saddle", C.c_int),
        ("internal_project_pressure", C.c_int),
        ("sc_gmres_restart", C.c_int),
        ("sc_lin_maxit_saddle", C.c_int),
        ("sc_lin_red_factor_saddle", C.c_double),
        ("sc_lin_res_norm_min_saddle", C.c_double),
        ("sc_div_factor", C.c_double),
        ("sc_minit", C.c_int),
        ("sc_amg_prec_it", C.c_int),
        ("precision", C.c_int),
        ("bf16_refine_steps", C.c_int),
        ("use_graph", C.c_int),
        ("shared_block", C.c_int),
        ("patch_inverse", C.c_int),
    ]


class PmgLevelDesc(C.Structure):
    _fields_ = [
        ("n_udof", C.c_int), ("n_active", C.c_int), ("hanging_bound", C.c_int),
        ("n_dirichlet", C.c_int), ("n_pdof", C.c_int),
        ("a_rowptr", c_int_p), ("a_kcol", c_int_p), ("a", c_double_p * 9), ("a_nnz", C.c_int),
        ("b_rowptr", c_int_p), ("b_kcol", c_int_p), ("b", c_double_p * 3), ("b_nnz", C.c_int),
        ("bt_rowptr", c_int_p), ("bt_kcol", c_int_p), ("bt", c_double_p * 3), ("bt_nnz", C.c_int),
        ("n_cells", C.c_int), ("n_ulocal", C.c_int), ("n_plocal", C.c_int),
        ("u_begin", c_int_p), ("u_global", c_int_p), ("p_begin", c_int_p), ("p_global", c_int_p),
        ("n_colors", C.c_int), ("ptr_colors", c_int_p), ("cell_reorder", c_int_p),
        ("parent", c_int_p), ("child", c_int_p), ("n_children", C.c_int),
        ("qq_u", c_double_p), ("qq_p", c_double_p),
    ]


class PmgInstanceDesc(C.Structure):
    _fields_ = [("n_levels", C.c_int), ("levels", C.POINTER(PmgLevelDesc)), ("project_pressure", C.c_int)]


class PmgKstat(C.Structure):
    _fields_ = [("ms", C.c_double), ("calls", C.c_longlong), ("bytes", C.c_double), ("flops", C.c_double)]


KSTATS = ["defect", "matvec", "vanka_res", "vanka_solve", "transfer", "coarse_solve", "patch_lu", "coarse_lu"]


class PmgTimers(C.Structure):
    _fields_ = [(k, PmgKstat) for k in KSTATS] + [("sweep_ms", C.c_double), ("sweep_calls", C.c_longlong),
                                                   ("vanka_res_finest", PmgKstat)]


class PmgpConfig(C.Structure):
    _fields_ = [
        ("cx", C.c_int), ("cy", C.c_int), ("cz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("n_levels", C.c_int), ("nu", C.c_double), ("dirichlet_faces", C.c_int),
        ("perturb", C.c_double), ("seed_perturb", C.c_ulonglong),
        ("conv_kind", C.c_int), ("conv_const", C.c_double * 3), ("seed_conv", C.c_ulonglong),
        ("conv_max", C.c_double), ("rhs_kind", C.c_int), ("seed_rhs", C.c_ulonglong),
        ("exact_kind", C.c_int), ("n_modes", C.c_int), ("mass", C.c_double),
    ]


CONV_NONE, CONV_FOURIER, CONV_CONST, CONV_CHANNEL = 0, 1, 2, 3
RHS_ZERO, RHS_FOURIER, RHS_EXACT = 0, 1, 2
EXACT_NONE, EXACT_QUADLIN = 0, 1

# every function declared in include/pmg.h, with its ctypes signature
PMG_FUNCS = {
    "pmg_params_default": (None, [C.POINTER(PmgParams)]),
    "pmg_status_string": (C.c_char_p, [C.c_int]),
    "pmg_ctx_create": (C.c_int, [C.c_int, C.POINTER(PmgParams), C.POINTER(C.c_void_p)]),
    "pmg_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pmg_last_error_string": (C.c_char_p, [C.c_void_p]),
    "pmg_set_params": (C.c_int, [C.c_void_p, C.POINTER(PmgParams)]),
    "pmg_level_add": (C.c_int, [C.c_void_p, C.POINTER(PmgLevelDesc), c_int_p]),
    "pmg_level_update_values": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PmgLevelDesc)]),
    "pmg_setup": (C.c_int, [C.c_void_p]),
    "pmg_refactor": (C.c_int, [C.c_void_p]),
    "pmg_level_info": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p]),
    "pmg_level_form": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "pmg_defect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, c_double_p, C.c_int]),
    "pmg_matvec": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_defect_nse4": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_vanka_smooth": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "pmg_vanka_smooth_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_restrict_defect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_prolongate": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_restrict_function": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_restrict_to_all_grids": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_int]),
    "pmg_correct_nodes": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "pmg_set_dirichlet": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_correct_defect": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "pmg_update": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_coarse_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_vcycle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_fgmres_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, c_int_p, c_double_p, C.c_int]),
    "pmg_fgmres_history": (C.c_int, [C.c_void_p, c_double_p, C.c_int]),
    "pmg_batch_create": (C.c_int, [C.c_int, C.POINTER(PmgParams), C.c_int, C.POINTER(PmgInstanceDesc),
                                   C.POINTER(C.c_void_p)]),
    "pmg_batch_fgmres_solve": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), c_int_p,
                                         c_double_p, c_int_p, C.c_int]),
    "pmg_batch_size": (C.c_int, [C.c_void_p]),
    "pmg_level_geometry": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "pmg_eval_velocity": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "pmg_assemble_oseen": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int]),
    "pmg_theta_rhs": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                C.c_void_p, C.c_int]),
    "pmg_color_cells": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, C.c_int, c_int_p, c_int_p, c_int_p]),
    "pmg_sequential_schedule": (C.c_int, [C.c_int, c_int_p, c_int_p, C.c_int, C.c_int, c_int_p, c_int_p, c_int_p,
                                          c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p]),
    "pmg_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "pmg_level_tuning": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "pmg_set_level_tuning": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "pmg_nodal_schedule": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                     c_int_p, c_int_p]),
    "pmg_level_read": (C.c_int, [C.c_char_p, C.POINTER(PmgLevelDesc), C.POINTER(C.c_void_p)]),
    "pmg_level_read_free": (C.c_int, [C.c_void_p]),
    "pmg_level_write": (C.c_int, [C.c_char_p, C.POINTER(PmgLevelDesc)]),
    "pmg_vector_read": (C.c_int, [C.c_char_p, c_double_p, C.c_long]),
    "pmg_vector_write": (C.c_int, [C.c_char_p, c_double_p, C.c_long]),
    "pmg_device_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "pmg_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pmg_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "pmg_synchronize": (C.c_int, [C.c_void_p]),
    "pmg_timers_get": (C.c_int, [C.c_void_p, C.POINTER(PmgTimers)]),
    "pmg_timers_reset": (C.c_int, [C.c_void_p]),
    "pmg_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
}

PROBLEM_FUNCS = {
    "pmgp_config_default": (None, [C.POINTER(PmgpConfig)]),
    "pmgp_create": (C.c_int, [C.POINTER(PmgpConfig), C.POINTER(C.c_void_p)]),
    "pmgp_destroy": (None, [C.c_void_p]),
    "pmgp_n_levels": (C.c_int, [C.c_void_p]),
    "pmgp_level_desc": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PmgLevelDesc)]),
    "pmgp_n_dof": (C.c_int, [C.c_void_p, C.c_int]),
    "pmgp_rhs": (c_double_p, [C.c_void_p, C.c_int]),
    "pmgp_exact": (c_double_p, [C.c_void_p, C.c_int]),
    "pmgp_dof_coords": (c_double_p, [C.c_void_p, C.c_int]),
    "pmgp_cell_centers": (c_double_p, [C.c_void_p, C.c_int]),
    "pmgp_conv_scale": (C.c_double, [C.c_void_p]),
    "pmgp_level_dims": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p]),
    "pmgp_cell_vertices": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "pmgp_conv_quad": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
}


class PmgError(RuntimeError):
    def __init__(self, status: int, msg: str = ""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}")


def _bind(lib: C.CDLL, table: dict) -> C.CDLL:
    for name, (res, args) in table.items():
        fn = getattr(lib, name)  # AttributeError if a declared symbol is missing
        fn.restype = res
        fn.argtypes = args
    return lib


_cache: dict[str, C.CDLL] = {}


def _load(name: str, table: dict) -> C.CDLL:
    if name in _cache:
        return _cache[name]
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(
            f"{path} is missing: build it with `python -m to_merge_master_parmoon_amd.build` "
            "(there is no fallback implementation)")
    lib = _bind(C.CDLL(str(path)), table)
    _cache[name] = lib
    return lib


def load_pmg() -> C.CDLL:
    """The HIP library (libpmg.so).  Loading does not touch the GPU."""
    return _load("libpmg.so", PMG_FUNCS)


def load_problem() -> C.CDLL:
    return _load("libpmg_problem.so", PROBLEM_FUNCS)


def default_params(**kw) -> PmgParams:
    p = PmgParams()
    load_pmg().pmg_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p
