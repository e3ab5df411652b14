"""ctypes binding of libemagddm_fetidp.so (the C ABI in include/emagddm_fetidp.h).

The shared library is built in-tree by ``emagddm_amd/csrc/Makefile``.  There is
no fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemagddm_fetidp.so")

SUCCESS, NUMERICAL_ERROR, DIVERGED, MAX_ITER_REACHED, INVALID_INPUT = range(5)
STATUS_NAMES = {0: "SUCCESS", 1: "NUMERICAL_ERROR", 2: "DIVERGED", 3: "MAX_ITER_REACHED",
                4: "INVALID_INPUT"}
FP64, FP32_REFINE = 0, 1


class Config(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_iterations", C.c_int32),
                ("tolerance", C.c_double), ("dual_tolerance", C.c_double),
                ("max_refinements", C.c_int32), ("verbose", C.c_int32),
                ("inner_tolerance", C.c_double), ("graphs", C.c_int32), ("precond_precision", C.c_int32),
                ("scaling", C.c_int32), ("side_stream", C.c_int32), ("tile_waves", C.c_int32),
                ("coarse_solve", C.c_int32), ("coarse_group", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("refinements", C.c_int32),
                ("fell_back", C.c_int32), ("residual_norm", C.c_double),
                ("dual_residual", C.c_double), ("initial_dual_residual", C.c_double),
                ("factor_ms", C.c_double), ("solve_ms", C.c_double), ("symbolic_ms", C.c_double),
                ("guess_ms", C.c_double), ("learned_guess", C.c_int32), ("precond_fp64", C.c_int32),
                ("deluxe", C.c_int32), ("reserved", C.c_int32)]


EXPORTS = [
    "fetidp_config_default", "fetidp_create", "fetidp_set_config", "fetidp_get_config",
    "fetidp_setup", "fetidp_set_materials",
    "fetidp_set_rhs_tet_J", "fetidp_set_rhs_free", "fetidp_factor", "fetidp_solve",
    "fetidp_get_sizes", "fetidp_get_choices", "fetidp_get_topology", "fetidp_get_global_nnz",
    "fetidp_get_global_system", "fetidp_apply_dual_operator", "fetidp_apply_preconditioner",
    "fetidp_get_dual_rhs", "fetidp_get_lambda", "fetidp_get_history",
    "fetidp_get_refinement_history",
    "fetidp_get_coarse_matrix", "fetidp_get_profile", "fetidp_reset_profile",
    "fetidp_last_error", "fetidp_destroy", "emag_nedelec_integrate", "emag_tile_cholesky_test",
    "emag_mfma_layout_probe", "fetidp_set_lambda0_model", "fetidp_get_dual_graph",
    "fetidp_get_lambda_features", "fetidp_autotune",
    "emag_mpnn_create", "emag_mpnn_destroy", "emag_mpnn_last_error", "emag_mpnn_num_params",
    "emag_mpnn_get_params", "emag_mpnn_set_params", "emag_mpnn_set_graph", "emag_mpnn_forward",
    "emag_mpnn_gradient", "emag_mpnn_train_step", "emag_mpnn_mfma_probe",
    "emag_problem_write", "emag_problem_read", "emag_problem_free", "emag_problem_last_error",
    "emag_assembly_create", "emag_assembly_sizes", "emag_assembly_pattern", "emag_assemble",
    "emag_assembly_last_error", "emag_assembly_destroy", "emag_csr_residual",
    "emag_csr_residual_noalloc", "emag_tree_gauge",
]
HEADERS = ["emagddm_fetidp.h", "emagddm_mpnn.h", "emagddm_problem.h", "emagddm_assembly.h"]


class Problem(C.Structure):
    """struct emag_problem (include/emagddm_problem.h)."""
    _fields_ = [("n_nodes", C.c_int64), ("n_tets", C.c_int64), ("n_pec", C.c_int64),
                ("n_mat", C.c_int32), ("n_sub", C.c_int32),
                ("xyz", C.POINTER(C.c_double)), ("tet2node", C.POINTER(C.c_int32)),
                ("tet_material", C.POINTER(C.c_int32)), ("tet_subdomain", C.POINTER(C.c_int32)),
                ("nu", C.POINTER(C.c_double)), ("kappa", C.POINTER(C.c_double)),
                ("pec_edge_nodes", C.POINTER(C.c_int32)), ("jq", C.POINTER(C.c_double))]

_lib = None


def load():
    """Load the in-tree library once; raise loudly if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C emagddm_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    lp = C.POINTER(C.c_int64)
    lib.fetidp_config_default.argtypes = [C.POINTER(Config)]
    lib.fetidp_config_default.restype = None
    lib.fetidp_create.argtypes = [C.POINTER(Config)]
    lib.fetidp_create.restype = P
    lib.fetidp_set_config.argtypes = [P, C.POINTER(Config)]
    lib.fetidp_get_config.argtypes = [P, C.POINTER(Config)]
    lib.fetidp_setup.argtypes = [P, C.c_int64, dp, C.c_int64, ip, ip, C.c_int32, dp, dp, ip,
                                 C.c_int64, ip]
    lib.fetidp_set_materials.argtypes = [P, C.c_int32, dp, dp]
    lib.fetidp_set_rhs_tet_J.argtypes = [P, dp]
    lib.fetidp_set_rhs_free.argtypes = [P, dp]
    lib.fetidp_factor.argtypes = [P]
    lib.fetidp_solve.argtypes = [P, dp, dp, C.POINTER(Stats)]
    lib.fetidp_get_sizes.argtypes = [P, lp]
    lib.fetidp_get_choices.argtypes = [P, ip, ip, dp]
    lib.fetidp_get_topology.argtypes = [P, ip, ip, ip, ip]
    lib.fetidp_get_global_nnz.argtypes = [P]
    lib.fetidp_get_global_nnz.restype = C.c_int64
    lib.fetidp_get_global_system.argtypes = [P, ip, ip, dp, dp]
    lib.fetidp_apply_dual_operator.argtypes = [P, dp, dp]
    lib.fetidp_apply_preconditioner.argtypes = [P, dp, dp]
    lib.fetidp_get_dual_rhs.argtypes = [P, dp]
    lib.fetidp_get_lambda.argtypes = [P, dp]
    lib.fetidp_get_history.argtypes = [P, dp, C.c_int32]
    lib.fetidp_get_refinement_history.argtypes = [P, dp, C.c_int32]
    lib.fetidp_get_coarse_matrix.argtypes = [P, dp]
    lib.fetidp_get_profile.argtypes = [P, dp, C.c_int32]
    lib.fetidp_reset_profile.argtypes = [P]
    lib.fetidp_reset_profile.restype = None
    lib.fetidp_last_error.argtypes = [P]
    lib.fetidp_last_error.restype = C.c_char_p
    lib.fetidp_destroy.argtypes = [P]
    lib.fetidp_destroy.restype = None
    lib.emag_nedelec_integrate.argtypes = [C.c_int64, P, P, P, P, P, P, P, P, P, P]
    lib.emag_tile_cholesky_test.argtypes = [C.c_int32, C.c_int64, C.c_int64, dp, dp]
    lib.emag_mfma_layout_probe.argtypes = [C.c_int32, dp]
    fp = C.POINTER(C.c_float)
    lib.fetidp_set_lambda0_model.argtypes = [P, P]
    lib.fetidp_get_dual_graph.argtypes = [P, ip, ip, fp]
    lib.fetidp_get_dual_graph.restype = C.c_int64
    lib.fetidp_get_lambda_features.argtypes = [P, fp, dp]
    lib.fetidp_autotune.argtypes = [P, C.c_int32, dp, C.c_int32]
    lib.emag_mpnn_create.argtypes = [C.c_uint64]
    lib.emag_mpnn_create.restype = P
    lib.emag_mpnn_destroy.argtypes = [P]
    lib.emag_mpnn_destroy.restype = None
    lib.emag_mpnn_last_error.argtypes = [P]
    lib.emag_mpnn_last_error.restype = C.c_char_p
    lib.emag_mpnn_num_params.argtypes = []
    lib.emag_mpnn_num_params.restype = C.c_int64
    lib.emag_mpnn_get_params.argtypes = [P, fp]
    lib.emag_mpnn_set_params.argtypes = [P, fp]
    lib.emag_mpnn_set_graph.argtypes = [P, C.c_int64, ip, ip, fp]
    lib.emag_mpnn_forward.argtypes = [P, C.c_int64, P, P, P]
    lib.emag_mpnn_gradient.argtypes = [P, C.c_int64, P, P, P, fp, dp, P]
    lib.emag_mpnn_train_step.argtypes = [P, C.c_int64, P, P, P, C.c_float, dp, P]
    lib.emag_mpnn_mfma_probe.argtypes = [fp, fp, fp]
    lib.emag_problem_write.argtypes = [C.c_char_p, C.POINTER(Problem)]
    lib.emag_problem_read.argtypes = [C.c_char_p, C.POINTER(Problem)]
    lib.emag_problem_free.argtypes = [C.POINTER(Problem)]
    lib.emag_problem_free.restype = None
    lib.emag_problem_last_error.argtypes = []
    lib.emag_problem_last_error.restype = C.c_char_p
    lib.emag_assembly_create.argtypes = [C.c_int64, C.c_int64, ip, C.c_int64, ip]
    lib.emag_assembly_create.restype = P
    lib.emag_assembly_sizes.argtypes = [P, lp, lp]
    lib.emag_assembly_pattern.argtypes = [P, ip, ip, ip, C.POINTER(C.c_int8)]
    lib.emag_assemble.argtypes = [P, P, P, P, P, P, P, P, P, P, P]
    lib.emag_assembly_last_error.argtypes = [P]
    lib.emag_assembly_last_error.restype = C.c_char_p
    lib.emag_assembly_destroy.argtypes = [P]
    lib.emag_assembly_destroy.restype = None
    lib.emag_csr_residual.argtypes = [C.c_int64, P, P, P, P, P, P, dp, P]
    lib.emag_csr_residual_noalloc.argtypes = [C.c_int64, P, P, P, P, P, P, dp, P]
    lib.emag_tree_gauge.argtypes = [C.c_int64, C.c_int64, ip, C.c_int64, ip, ip]
    lib.emag_tree_gauge.restype = C.c_int64
    for name in EXPORTS:
        fn = getattr(lib, name)
        if fn.restype is C.c_int:  # default int return -> status
            fn.restype = C.c_int32
    _lib = lib
    return lib
