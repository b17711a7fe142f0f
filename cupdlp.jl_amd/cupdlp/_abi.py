"""ctypes mirror of include/pdlp_hip.h (the C ABI of libpdlp_hip.so).

Every structure here has the same field order and types as the header; the
library reports its own struct sizes through ``pdlp_abi_layout`` and
``check_layout`` compares them at load time.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libpdlp_hip.so")


class pdlp_lp(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("m", C.c_int64), ("num_eq", C.c_int64), ("nnz", C.c_int64),
        ("colptr", C.POINTER(C.c_int64)), ("rowval", C.POINTER(C.c_int64)),
        ("nzval", C.POINTER(C.c_double)),
        ("c", C.POINTER(C.c_double)), ("lb", C.POINTER(C.c_double)),
        ("ub", C.POINTER(C.c_double)), ("b", C.POINTER(C.c_double)),
        ("c0", C.c_double),
    ]


class pdlp_mps_problem(C.Structure):
    _fields_ = [
        ("lp", pdlp_lp), ("num_file_variables", C.c_int64), ("num_slack_variables", C.c_int64),
        ("quadratic_entries", C.c_int64), ("owner", C.c_void_p),
    ]


class pdlp_termination_criteria(C.Structure):
    _fields_ = [
        ("optimality_norm", C.c_int32), ("pad0_", C.c_int32),
        ("eps_optimal_absolute", C.c_double), ("eps_optimal_relative", C.c_double),
        ("eps_primal_infeasible", C.c_double), ("eps_dual_infeasible", C.c_double),
        ("time_sec_limit", C.c_double), ("iteration_limit", C.c_int32), ("pad1_", C.c_int32),
        ("kkt_matrix_pass_limit", C.c_double),
    ]


class pdlp_restart_params(C.Structure):
    _fields_ = [
        ("restart_scheme", C.c_int32), ("restart_to_current_metric", C.c_int32),
        ("restart_frequency_if_fixed", C.c_int64),
        ("artificial_restart_threshold", C.c_double),
        ("sufficient_reduction_for_restart", C.c_double),
        ("necessary_reduction_for_restart", C.c_double),
        ("primal_weight_update_smoothing", C.c_double),
    ]


class pdlp_params(C.Structure):
    _fields_ = [
        ("l_inf_ruiz_iterations", C.c_int64),
        ("l2_norm_rescaling", C.c_int32),
        ("has_pock_chambolle_alpha", C.c_int32),
        ("pock_chambolle_alpha", C.c_double),
        ("primal_importance", C.c_double),
        ("scale_invariant_initial_primal_weight", C.c_int32),
        ("pad0_", C.c_int32),
        ("verbosity", C.c_int64),
        ("record_iteration_stats", C.c_int32),
        ("termination_evaluation_frequency", C.c_int32),
        ("termination_criteria", pdlp_termination_criteria),
        ("restart_params", pdlp_restart_params),
        ("step_size_policy", C.c_int32),
        ("pad1_", C.c_int32),
        ("reduction_exponent", C.c_double),
        ("growth_exponent", C.c_double),
        ("precision", C.c_int32),
        ("refine_max_outer", C.c_int32),
        ("refine_inner_tolerance", C.c_double),
        ("refine_max_scale", C.c_double),
        ("device_preconditioning", C.c_int32),
        ("refine_inner_iteration_limit", C.c_int32),
        ("batch_groups", C.c_int32),
        ("refine_early_fallback", C.c_int32),
        ("debug_numerical_error_attempt", C.c_int32),
        ("refine_warm_start", C.c_int32),
        ("lane_budget", C.c_int32),
        ("slice_unroll", C.c_int32),
        ("slice_waves", C.c_int32),
        ("refine_continue_progress", C.c_int32),
        ("primal_occupancy", C.c_int32),
        ("dual_occupancy", C.c_int32),
    ]


class pdlp_convergence_information(C.Structure):
    _fields_ = [("candidate_type", C.c_int32), ("pad_", C.c_int32)] + [
        (name, C.c_double) for name in (
            "primal_objective", "dual_objective", "corrected_dual_objective",
            "l_inf_primal_residual", "l2_primal_residual", "l_inf_dual_residual",
            "l2_dual_residual", "relative_l_inf_primal_residual",
            "relative_l2_primal_residual", "relative_l_inf_dual_residual",
            "relative_l2_dual_residual", "relative_optimality_gap",
            "l_inf_primal_variable", "l2_primal_variable", "l_inf_dual_variable",
            "l2_dual_variable")
    ]


class pdlp_infeasibility_information(C.Structure):
    _fields_ = [("candidate_type", C.c_int32), ("pad_", C.c_int32)] + [
        (name, C.c_double) for name in (
            "max_primal_ray_infeasibility", "primal_ray_linear_objective",
            "max_dual_ray_infeasibility", "dual_ray_objective")
    ]


class pdlp_iteration_stats(C.Structure):
    _fields_ = [
        ("iteration_number", C.c_int32),
        ("cumulative_rejected_steps", C.c_int32),
        ("convergence_information", pdlp_convergence_information),
        ("infeasibility_information", pdlp_infeasibility_information),
        ("cumulative_kkt_matrix_passes", C.c_double),
        ("cumulative_time_sec", C.c_double),
        ("restart_used", C.c_int32),
        ("pad_", C.c_int32),
        ("step_size", C.c_double),
        ("primal_weight", C.c_double),
        ("time_spent_doing_basic_algorithm", C.c_double),
        ("total_attempts", C.c_int64),
    ]


class pdlp_result(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("x_current", C.POINTER(C.c_double)), ("y_current", C.POINTER(C.c_double)),
        ("termination_reason", C.c_int32), ("iteration_count", C.c_int32),
        ("final_stats", pdlp_iteration_stats),
        ("solve_time_sec", C.c_double), ("preprocess_time_sec", C.c_double),
        ("total_attempts", C.c_int64),
        ("refinement_rounds", C.c_int32), ("fp64_fallback", C.c_int32),
    ]


class pdlp_kernel_profile(C.Structure):
    _fields_ = [("primal_launches", C.c_int64), ("dual_launches", C.c_int64),
                ("primal_ms", C.c_double), ("dual_ms", C.c_double),
                ("primal_bytes", C.c_double), ("dual_bytes", C.c_double),
                ("fallback_primal_launches", C.c_int64), ("fallback_dual_launches", C.c_int64),
                ("fallback_primal_ms", C.c_double), ("fallback_dual_ms", C.c_double),
                ("fallback_primal_bytes", C.c_double), ("fallback_dual_bytes", C.c_double)]


class pdlp_tuning(C.Structure):
    _fields_ = [("lane_budget", C.c_int32), ("slice_unroll", C.c_int32),
                ("slice_waves", C.c_int32), ("candidates", C.c_int32),
                ("kernel_us_per_iteration", C.c_double),
                ("default_us_per_iteration", C.c_double), ("seconds", C.c_double),
                ("primal_occupancy", C.c_int32), ("dual_occupancy", C.c_int32)]


STATS_CB = C.CFUNCTYPE(None, C.POINTER(pdlp_iteration_stats), C.c_void_p)

# Order of the sizes reported by pdlp_abi_layout().
LAYOUT_ORDER = (pdlp_lp, pdlp_termination_criteria, pdlp_restart_params, pdlp_params,
                pdlp_convergence_information, pdlp_infeasibility_information,
                pdlp_iteration_stats, pdlp_result)

# Every function the header declares (checked by the CPU test suite).
EXPORTED = (
    "pdlp_ctx_create", "pdlp_ctx_destroy", "pdlp_last_error", "pdlp_default_params",
    "pdlp_solve", "pdlp_batch_solve", "pdlp_serve_open", "pdlp_serve_submit",
    "pdlp_serve_poll", "pdlp_serve_result", "pdlp_serve_close", "pdlp_abi_version",
    "pdlp_abi_layout", "pdlp_set_profiling", "pdlp_get_profile",
    "pdlp_batch_load", "pdlp_batch_run", "pdlp_batch_free",
    "pdlp_read_mps", "pdlp_mps_free", "pdlp_serve_submit_limits", "pdlp_serve_cancel",
    "pdlp_solve_noalloc", "pdlp_tune",
)


def declare(lib):
    """Attach argtypes/restypes for the exported functions."""
    lib.pdlp_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.pdlp_ctx_create.restype = C.c_int
    lib.pdlp_ctx_destroy.argtypes = [C.c_void_p]
    lib.pdlp_ctx_destroy.restype = None
    lib.pdlp_last_error.argtypes = [C.c_void_p]
    lib.pdlp_last_error.restype = C.c_char_p
    lib.pdlp_default_params.argtypes = [C.POINTER(pdlp_params)]
    lib.pdlp_default_params.restype = None
    lib.pdlp_solve.argtypes = [C.c_void_p, C.POINTER(pdlp_lp), C.POINTER(pdlp_params),
                               C.POINTER(pdlp_result), STATS_CB, C.c_void_p]
    lib.pdlp_solve.restype = C.c_int
    lib.pdlp_solve_noalloc.argtypes = lib.pdlp_solve.argtypes
    lib.pdlp_solve_noalloc.restype = C.c_int
    lib.pdlp_tune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(pdlp_lp), C.POINTER(pdlp_params),
                              C.c_int32, C.POINTER(pdlp_tuning)]
    lib.pdlp_tune.restype = C.c_int
    lib.pdlp_batch_solve.argtypes = [C.c_void_p, C.c_int32, C.POINTER(pdlp_lp),
                                     C.POINTER(pdlp_params), C.POINTER(pdlp_result)]
    lib.pdlp_batch_solve.restype = C.c_int
    lib.pdlp_batch_load.argtypes = [C.c_void_p, C.c_int32, C.POINTER(pdlp_lp),
                                    C.POINTER(pdlp_params), C.POINTER(C.c_void_p)]
    lib.pdlp_batch_load.restype = C.c_int
    lib.pdlp_batch_run.argtypes = [C.c_void_p, C.POINTER(pdlp_result)]
    lib.pdlp_batch_run.restype = C.c_int
    lib.pdlp_batch_free.argtypes = [C.c_void_p]
    lib.pdlp_batch_free.restype = None
    lib.pdlp_read_mps.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.POINTER(pdlp_mps_problem)),
                                  C.c_char_p, C.c_int64]
    lib.pdlp_read_mps.restype = C.c_int
    lib.pdlp_mps_free.argtypes = [C.POINTER(pdlp_mps_problem)]
    lib.pdlp_mps_free.restype = None
    lib.pdlp_serve_open.argtypes = [C.c_void_p, C.POINTER(pdlp_params), C.c_double, C.c_int32,
                                    C.POINTER(C.c_void_p)]
    lib.pdlp_serve_open.restype = C.c_int
    lib.pdlp_serve_submit.argtypes = [C.c_void_p, C.POINTER(pdlp_lp), C.c_double,
                                      C.POINTER(C.c_int64)]
    lib.pdlp_serve_submit.restype = C.c_int
    lib.pdlp_serve_submit_limits.argtypes = [C.c_void_p, C.POINTER(pdlp_lp), C.c_double,
                                             C.c_double, C.c_int32, C.POINTER(C.c_int64)]
    lib.pdlp_serve_submit_limits.restype = C.c_int
    lib.pdlp_serve_cancel.argtypes = [C.c_void_p, C.c_int64]
    lib.pdlp_serve_cancel.restype = C.c_int
    lib.pdlp_serve_poll.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32,
                                    C.POINTER(C.c_int32)]
    lib.pdlp_serve_poll.restype = C.c_int
    lib.pdlp_serve_result.argtypes = [C.c_void_p, C.c_int64, C.POINTER(pdlp_result)]
    lib.pdlp_serve_result.restype = C.c_int
    lib.pdlp_serve_close.argtypes = [C.c_void_p]
    lib.pdlp_serve_close.restype = C.c_int
    lib.pdlp_abi_version.argtypes = []
    lib.pdlp_abi_version.restype = C.c_int
    lib.pdlp_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    lib.pdlp_set_profiling.restype = C.c_int
    lib.pdlp_get_profile.argtypes = [C.c_void_p, C.POINTER(pdlp_kernel_profile)]
    lib.pdlp_get_profile.restype = C.c_int
    lib.pdlp_abi_layout.argtypes = [C.POINTER(C.c_int64), C.c_int32]
    lib.pdlp_abi_layout.restype = C.c_int32
    return lib


def check_layout(lib):
    """Raise if the library's struct sizes differ from this mirror's."""
    buf = (C.c_int64 * 16)()
    cnt = lib.pdlp_abi_layout(buf, 16)
    want = [C.sizeof(t) for t in LAYOUT_ORDER]
    got = list(buf[:cnt])
    if got != want:
        raise RuntimeError(f"pdlp ABI layout mismatch: library {got} vs ctypes {want}")


_LIB = None


def load(path=None):
    """Load libpdlp_hip.so (built in-tree); raises if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libpdlp_hip.so not found at {p}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = declare(C.CDLL(p))
    check_layout(lib)
    if path is None:
        _LIB = lib
    return lib
