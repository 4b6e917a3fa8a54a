"""ctypes binding of libchfsi.so (the C ABI declared in include/chfsi.h).

The library is the only compute path: if it is missing this module raises, it
never falls back to CPU code.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchfsi.so")

c_double_p = ct.POINTER(ct.c_double)
c_float_p = ct.POINTER(ct.c_float)
c_int64_p = ct.POINTER(ct.c_int64)
c_int32_p = ct.POINTER(ct.c_int32)
c_int8_p = ct.POINTER(ct.c_int8)

STATUS = {0: "CHFSI_OK", 1: "CHFSI_ERR_ARG", 2: "CHFSI_ERR_HIP", 3: "CHFSI_ERR_BLAS",
          4: "CHFSI_ERR_NOT_SPD", 5: "CHFSI_ERR_EIG", 6: "CHFSI_ERR_STATE",
          7: "CHFSI_ERR_NOMEM", 8: "CHFSI_ERR_NAN"}

FP64, FP32, BF16 = 0, 1, 2


class ChfsiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class OperatorDesc(ct.Structure):
    _fields_ = [
        ("n_dofs", ct.c_int64), ("n_cells", ct.c_int32), ("nodes_per_cell", ct.c_int32),
        ("quad_per_cell", ct.c_int32), ("cell_map", c_int64_p), ("cell_color", c_int8_p),
        ("sqrt_mass", c_double_p), ("inv_sqrt_mass", c_double_p),
        ("shape_values", c_double_p), ("stiffness", c_double_p),
        ("n_proj", ct.c_int32), ("n_pairs", ct.c_int32), ("proj_per_pair", ct.c_int32),
        ("pair_cell", c_int32_p), ("pair_proj0", c_int32_p), ("pair_C", c_double_p),
        ("proj_V", c_double_p),
        ("n_constrained", ct.c_int64), ("constraint_rows", c_int64_p),
        ("constraint_ptr", c_int64_p), ("constraint_cols", c_int64_p),
        ("constraint_vals", c_double_p),
        ("jxw", c_double_p),
        ("n_kpoints", ct.c_int32), ("kpoints", c_double_p), ("cell_mass", c_double_p),
        ("cell_grad", c_double_p), ("pair_C_kpoints", c_double_p),
        ("shape_gradients", c_double_p),
        ("cell_stiffness", c_double_p), ("inverse_jacobians", c_double_p),
        ("jacobian_kind", ct.c_int32),
    ]


class KBatch(ct.Structure):
    _fields_ = [
        ("n", ct.c_int32), ("kpoint", c_int32_p), ("X", ct.POINTER(ct.c_void_p)),
        ("ld", c_int64_p), ("n_vec", c_int32_p), ("a", c_double_p), ("b", c_double_p),
        ("a0", c_double_p), ("degree", ct.c_int32), ("precision", ct.c_int32),
    ]


class Params(ct.Structure):
    _fields_ = [
        ("cheby_block_size", ct.c_int32), ("cheby_degree", ct.c_int32),
        ("lanczos_steps", ct.c_int32), ("reproducible", ct.c_int32),
        ("reuse_lanczos_upper", ct.c_int32), ("first_scf_degree_factor", ct.c_double),
        ("filter_precision", ct.c_int32), ("use_shep", ct.c_int32),
        ("fp64_tail_degrees", ct.c_int32),
        ("mixed_prec_overlap", ct.c_int32), ("mixed_prec_cgs_rotation", ct.c_int32),
        ("mixed_prec_xthx", ct.c_int32), ("mixed_prec_rr_rotation", ct.c_int32),
        ("mixed_prec_xthx_spectrum_split", ct.c_int32),
        ("filter_columns", ct.c_int32),
        ("mixed_prec_switch_residual", ct.c_double),
        ("mixed_prec_defect_filter", ct.c_int32),
        ("is_pseudopotential", ct.c_int32), ("lanczos_full_residual_margin", ct.c_int32),
        ("mixed_prec_fp64_top_columns", ct.c_int32), ("dense_rocblas", ct.c_int32),
    ]


# name -> (restype, argtypes)
_VP = ct.c_void_p
SIGNATURES = {
    "chfsi_last_error": (ct.c_char_p, []),
    "chfsi_default_params": (None, [ct.POINTER(Params)]),
    "chfsi_chebyshev_degree": (ct.c_int, [ct.POINTER(Params), ct.c_double, ct.c_int32,
                                          c_int32_p]),
    "chfsi_operator_create": (ct.c_int, [ct.POINTER(OperatorDesc), ct.POINTER(_VP)]),
    "chfsi_operator_destroy": (None, [_VP]),
    "chfsi_operator_set_potential": (ct.c_int, [_VP, _VP, _VP]),
    "chfsi_operator_set_potential_gga": (ct.c_int, [_VP, _VP, _VP, _VP]),
    "chfsi_operator_set_cell_matrices": (ct.c_int, [_VP, _VP, _VP]),
    "chfsi_operator_get_cell_matrices": (ct.c_int, [_VP, _VP, _VP]),
    "chfsi_operator_set_hx_variant": (ct.c_int, [_VP, ct.c_int32, ct.c_int32]),
    "chfsi_operator_set_proj_variant": (ct.c_int, [_VP, ct.c_int32, ct.c_int32]),
    "chfsi_operator_autotune": (ct.c_int, [_VP, ct.c_int32, c_int32_p, c_int32_p,
                                           ct.POINTER(ct.c_float), _VP]),
    "chfsi_operator_sqrt_mass": (_VP, [_VP]),
    "chfsi_operator_inv_sqrt_mass": (_VP, [_VP]),
    "chfsi_operator_n_dofs": (ct.c_int64, [_VP]),
    "chfsi_hx": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                            ct.c_double, _VP]),
    "chfsi_hx_noalloc": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32,
                                    ct.c_int32, ct.c_double, _VP]),
    "chfsi_hx_fp32": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                 ct.c_float, _VP]),
    "chfsi_hx_bf16": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                 ct.c_float, _VP]),
    "chfsi_hx_cheby": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_density": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP, _VP, ct.c_int32, _VP]),
    "chfsi_density_complex": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP, _VP, ct.c_int32,
                                         _VP]),
    "chfsi_hx_adjoint_v": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32,
                                      ct.c_double, _VP, ct.c_int32, _VP]),
    "chfsi_chebyshev_filter_backward": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64,
                                                   ct.c_int32, ct.c_int32, ct.c_double,
                                                   ct.c_double, ct.c_double, ct.c_int32, _VP,
                                                   ct.c_int64, _VP, ct.c_int32, _VP]),
    "chfsi_constraints_distribute": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_constraints_slave_to_master": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_constraints_set_zero": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_chebyshev_filter": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                          ct.c_double, ct.c_double, ct.c_double, ct.c_int32, _VP]),
    "chfsi_lanczos_bounds": (ct.c_int, [_VP, _VP, ct.c_int32, ct.c_int32, c_double_p, c_double_p,
                                        _VP]),
    "chfsi_cholgs": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_rayleigh_ritz": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32, c_double_p,
                                       _VP]),
    "chfsi_residual_norms": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32, c_double_p,
                                        c_double_p, _VP]),
    "chfsi_solver_create": (ct.c_int, [ct.POINTER(Params), ct.c_double, ct.c_double, ct.c_double,
                                       ct.POINTER(_VP)]),
    "chfsi_solver_destroy": (None, [_VP]),
    "chfsi_solver_set_params": (ct.c_int, [_VP, ct.POINTER(Params)]),
    "chfsi_solver_reinit_bounds": (ct.c_int, [_VP, ct.c_double, ct.c_double, ct.c_double]),
    "chfsi_solver_get_bounds": (ct.c_int, [_VP, c_double_p, c_double_p, c_double_p]),
    "chfsi_solver_last_filter_precision": (ct.c_int, [_VP, c_int32_p]),
    "chfsi_solver_set_lanczos_start": (ct.c_int, [_VP, _VP, ct.c_int64]),
    "chfsi_solve": (ct.c_int, [_VP, _VP, _VP, ct.c_int64, ct.c_int32, c_double_p, c_double_p,
                               ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, c_double_p,
                               c_double_p, _VP]),
    # spectrum split (top N - n_core Ritz pairs)
    "chfsi_rayleigh_ritz_spectrum_split": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                                      _VP, ct.c_int64, ct.c_int32, c_double_p,
                                                      ct.c_int32, _VP]),
    # filter + CholGS passes without Rayleigh-Ritz (solveNoRR)
    "chfsi_solve_no_rr": (ct.c_int, [_VP, _VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                     ct.c_int32, ct.c_int32, _VP]),
    "chfsi_solve_spectrum_split": (ct.c_int, [_VP, _VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                              _VP, ct.c_int64, c_double_p, c_double_p,
                                              ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32,
                                              c_double_p, c_double_p, _VP]),
    # complex (Bloch k-point) form
    "chfsi_operator_set_kpoint": (ct.c_int, [_VP, ct.c_int32]),
    "chfsi_operator_n_kpoints": (ct.c_int32, [_VP]),
    "chfsi_hx_complex": (ct.c_int, [_VP, _VP, ct.c_int64, _VP, ct.c_int64, ct.c_int32,
                                    ct.c_int32, ct.c_double, ct.c_int32, _VP]),
    "chfsi_chebyshev_filter_complex": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                                  ct.c_double, ct.c_double, ct.c_double,
                                                  ct.c_int32, _VP]),
    "chfsi_batch_filter": (ct.c_int, [_VP, ct.POINTER(KBatch), _VP]),
    "chfsi_batch_filter_create": (ct.c_int, [_VP, ct.POINTER(KBatch), _VP, ct.POINTER(_VP)]),
    "chfsi_batch_filter_set_bounds": (ct.c_int, [_VP, c_double_p, c_double_p, c_double_p]),
    "chfsi_batch_filter_run": (ct.c_int, [_VP, _VP]),
    "chfsi_batch_filter_launch": (ct.c_int, [_VP, _VP]),
    "chfsi_batch_filter_destroy": (None, [_VP]),
    "chfsi_operator_set_hx_variant_complex": (ct.c_int, [_VP, ct.c_int32, ct.c_int32]),
    "chfsi_operator_set_proj_variant_complex": (ct.c_int, [_VP, ct.c_int32]),
    "chfsi_operator_autotune_complex": (ct.c_int, [_VP, ct.POINTER(KBatch), c_int32_p, c_int32_p,
                                                   ct.POINTER(ct.c_float), _VP]),
    "chfsi_lanczos_bounds_complex": (ct.c_int, [_VP, _VP, ct.c_int32, ct.c_int32, c_double_p,
                                                c_double_p, _VP]),
    "chfsi_cholgs_complex": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, _VP]),
    "chfsi_rayleigh_ritz_complex": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                               c_double_p, _VP]),
    "chfsi_residual_norms_complex": (ct.c_int, [_VP, _VP, ct.c_int64, ct.c_int32, ct.c_int32,
                                                c_double_p, c_double_p, _VP]),
    "chfsi_solve_complex": (ct.c_int, [_VP, _VP, _VP, ct.c_int64, ct.c_int32, c_double_p,
                                       c_double_p, ct.c_int32, ct.c_int32, ct.c_int32,
                                       ct.c_int32, c_double_p, c_double_p, _VP]),
}

_lib = None


def load() -> ct.CDLL:
    """Load libchfsi.so and bind every declared symbol (raises if anything is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m dftfe_amd.build`")
    lib = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().chfsi_last_error()
        raise ChfsiError(status, msg.decode() if msg else "")
