"""CPU checks of the C ABI library and the host-side mirror (no GPU calls)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import lp_cases
import oracle as O
from cupdlp import (Optimizer, PdhgParameters, TerminationReason as TR, _abi,
                   construct_termination_criteria, termination_reason_to_string)
from cupdlp.solver import to_c_lp, to_c_params

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "pdlp_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(pdlp_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.EXPORTED)


def test_abi_layout_matches_ctypes_mirror():
    lib = _abi.load()
    _abi.check_layout(lib)
    assert lib.pdlp_abi_version() == 2


def test_struct_field_order_is_the_references():
    """pdlp_termination_criteria follows src/termination.jl:9-77 and
    pdlp_params src/primal_dual_hybrid_gradient_gpu.jl:9-24 field for field
    (pad_ fields are C alignment; Union{Float64,Nothing} is has_ + value; the
    step-size policy variant is step_size_policy + AdaptiveStepsizeParams'
    two fields), then the extensions."""
    def names(t):
        return [f for f, _ in t._fields_ if not f.startswith("pad")]
    assert names(_abi.pdlp_termination_criteria) == [
        "optimality_norm", "eps_optimal_absolute", "eps_optimal_relative",
        "eps_primal_infeasible", "eps_dual_infeasible", "time_sec_limit", "iteration_limit",
        "kkt_matrix_pass_limit"]
    head = names(_abi.pdlp_params)[:15]
    assert head == [
        "l_inf_ruiz_iterations", "l2_norm_rescaling", "has_pock_chambolle_alpha",
        "pock_chambolle_alpha", "primal_importance", "scale_invariant_initial_primal_weight",
        "verbosity", "record_iteration_stats", "termination_evaluation_frequency",
        "termination_criteria", "restart_params", "step_size_policy", "reduction_exponent",
        "growth_exponent", "precision"]
    assert names(_abi.pdlp_restart_params) == [
        "restart_scheme", "restart_to_current_metric", "restart_frequency_if_fixed",
        "artificial_restart_threshold", "sufficient_reduction_for_restart",
        "necessary_reduction_for_restart", "primal_weight_update_smoothing"]


def test_default_params_match_reference_defaults():
    lib = _abi.load()
    p = _abi.pdlp_params()
    lib.pdlp_default_params(C.byref(p))
    mirror = to_c_params(PdhgParameters())
    for f, _ in _abi.pdlp_params._fields_:
        if f in ("termination_criteria", "restart_params", "reserved") or f.startswith("pad"):
            continue
        assert getattr(p, f) == getattr(mirror, f), f
    for sub in ("termination_criteria", "restart_params"):
        for f, _ in getattr(_abi, "pdlp_" + sub)._fields_:
            assert getattr(getattr(p, sub), f) == getattr(getattr(mirror, sub), f), (sub, f)
    # src/MOI_wrapper.jl:58-78
    assert p.l_inf_ruiz_iterations == 10 and p.pock_chambolle_alpha == 1.0
    assert p.termination_evaluation_frequency == 64
    assert p.restart_params.artificial_restart_threshold == 0.36
    assert p.termination_criteria.eps_optimal_absolute == 1e-6


def test_termination_strings():
    assert termination_reason_to_string(TR.TERMINATION_REASON_OPTIMAL) == "OPTIMAL"
    assert termination_reason_to_string(7) == "NUMERICAL_ERROR"


def test_lp_marshalling_is_one_based_csc():
    prob, _ = lp_cases.tiny_two_var()
    lp, keep = to_c_lp(prob)
    assert lp.n == 2 and lp.m == 2 and lp.num_eq == 1 and lp.nnz == 4
    assert [lp.colptr[i] for i in range(3)] == [1, 3, 5]
    assert [lp.rowval[i] for i in range(4)] == [1, 2, 1, 2]


def test_ctx_create_without_device_fails_cleanly():
    lib = _abi.load()
    h = C.c_void_p()
    try:
        n = C.c_int(0)
        has = C.CDLL("libamdhip64.so").hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        has = False
    if has:
        pytest.skip("a device is present")
    rc = lib.pdlp_ctx_create(0, C.byref(h))
    assert rc != 0
    assert lib.pdlp_last_error(None)


# ---- the MOI-style front end (src/MOI_wrapper.jl semantics), solved by the oracle

def _oracle_solve(params, problem):
    return O.solve(problem, params)


def test_optimizer_row_order_sense_and_duals():
    # max x + y  s.t.  x + 2y <= 4 (ge after negation), x - y == 1, 0 <= x, y <= 3
    opt = Optimizer(solve_fn=_oracle_solve)
    opt.silent = True
    tc = construct_termination_criteria(eps_optimal_absolute=1e-9, eps_optimal_relative=1e-9)
    opt.set_raw("termination_criteria", tc)
    x = opt.add_variable(0.0, math.inf)
    y = opt.add_variable(-math.inf, 3.0)
    c_le = opt.add_constraint({x: 1.0, y: 2.0}, "<=", 4.0)
    c_eq = opt.add_constraint({x: 1.0, y: -1.0}, "==", 1.0)
    opt.set_objective({x: 1.0, y: 1.0}, 0.5, sense="max")
    opt.optimize()
    assert opt.termination_status() == "OPTIMAL"
    assert opt.primal_status() == "FEASIBLE_POINT"
    # optimum: x - y = 1, x + 2y = 4 -> y = 1, x = 2, objective 3.5
    assert abs(opt.variable_primal(x) - 2.0) < 1e-6
    assert abs(opt.variable_primal(y) - 1.0) < 1e-6
    assert abs(opt.objective_value() - 3.5) < 1e-6
    assert opt.num_equalities == 1
    # equality row is stored first; the <= row reads dual_solution[num_eq + 0], negated
    out = opt.output
    assert opt.constraint_dual(c_eq) == out.dual_solution[0]
    assert opt.constraint_dual(c_le) == -out.dual_solution[1]
    assert opt.raw_status_string().endswith(": OPTIMAL")
    assert opt.solve_time_sec() >= 0.0


def test_optimizer_statuses_for_infeasible():
    opt = Optimizer(solve_fn=_oracle_solve)
    x = opt.add_variable(-math.inf, 0.0)
    opt.add_constraint({x: 1.0}, ">=", 1.0)
    opt.set_objective({x: 1.0})
    opt.optimize()
    assert opt.termination_status() == "INFEASIBLE"
    assert opt.primal_status() == "NO_SOLUTION"
    assert opt.dual_status() == "INFEASIBILITY_CERTIFICATE"


def test_optimizer_raw_attributes():
    opt = Optimizer(solve_fn=_oracle_solve)
    opt.set_raw("l_inf_ruiz_iterations", 3)
    assert opt.get_raw("l_inf_ruiz_iterations") == 3
    with pytest.raises(KeyError):
        opt.set_raw("no_such_field", 1)
    assert opt.result_count() == 0 and opt.termination_status() == "OPTIMIZE_NOT_CALLED"
