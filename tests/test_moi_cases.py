"""The MOI-style front end on known-answer models (tests/moi_cases.py), solved
by the oracle: every stage's statuses and answers are the hand-derived ones,
and HiGHS (scipy.optimize.linprog) agrees on the very problem the front end
built -- its status (optimal / infeasible / unbounded) and optimal objective.
Interval rows, vector rows, repeated terms, modifications, deletions, offsets
and MAX sense are the features of the MathOptInterface.Test functions the
reference runs (test/tmp.jl:28-67)."""
import math

import numpy as np
import pytest
from scipy.optimize import linprog

import moi_cases
import oracle as O
from cupdlp import Optimizer, construct_termination_criteria
from cupdlp.moi import InvalidIndex, ResultIndexBoundsError

TOL = 1e-6


def oracle_optimizer():
    def solve(params, problem):
        return O.solve(problem, params)
    opt = Optimizer(solve_fn=solve)
    opt.silent = True
    opt.set_raw("termination_criteria",
                construct_termination_criteria(eps_optimal_absolute=1e-9, eps_optimal_relative=1e-9))
    return opt


def highs(prob):
    """(status, objective of the min problem): 0 optimal, 2 infeasible, 3 unbounded."""
    A = prob.constraint_matrix.tocsr()
    k = prob.num_equalities
    bounds = [(None if not np.isfinite(lo) else lo, None if not np.isfinite(hi) else hi)
              for lo, hi in zip(prob.variable_lower_bound, prob.variable_upper_bound)]
    kw = {}
    if k > 0:
        kw["A_eq"], kw["b_eq"] = A[:k].toarray(), prob.right_hand_side[:k]
    if A.shape[0] > k:
        kw["A_ub"], kw["b_ub"] = -A[k:].toarray(), -prob.right_hand_side[k:]
    r = linprog(prob.objective_vector, bounds=bounds, method="highs", **kw)
    return r.status, (r.fun + prob.objective_constant if r.status == 0 else None)


@pytest.mark.parametrize("name", list(moi_cases.CASES))
def test_moi_case(name):
    stages = moi_cases.CASES[name](oracle_optimizer())
    for k, st in enumerate(stages):
        where = f"{name} stage {k}"
        assert st.termination == st.expected_status, where
        status, fun = highs(st.problem)
        assert status == {"OPTIMAL": 0, "INFEASIBLE": 2, "DUAL_INFEASIBLE": 3}[st.expected_status], where
        if st.expected_status == "OPTIMAL":
            assert st.primal_status == "FEASIBLE_POINT" and st.dual_status == "FEASIBLE_POINT"
            assert abs(st.objective - st.expected_objective) <= TOL * (1 + abs(st.expected_objective)), where
            assert abs(st.dual_objective - st.expected_objective) <= TOL * (1 + abs(st.expected_objective))
            # HiGHS solves the problem handed to the solver (a MAX objective negated)
            assert abs(fun - (-st.expected_objective if st.max_sense else st.expected_objective)) <= 1e-9
            for v, want in st.expected_primal.items():
                assert abs(st.primal[v] - want) <= TOL, (where, v)
            for c, want in st.expected_duals.items():
                assert abs(st.duals[c] - want) <= TOL, (where, c)
        elif st.expected_status == "INFEASIBLE":
            assert st.primal_status == "NO_SOLUTION"
            assert st.dual_status == "INFEASIBILITY_CERTIFICATE"
        else:
            assert st.primal_status == "INFEASIBILITY_CERTIFICATE"
            assert st.dual_status == "NO_SOLUTION"


def test_moi_vector_duals_and_values():
    st = moi_cases.conic_linear_vector_affine(oracle_optimizer())[0]
    # x, z basic: x's reduced cost gives the first row's dual -3, z's the second's -1;
    # y sits at its bound with multiplier -2 + 3 + 1 = 2
    np.testing.assert_allclose(st.vector_duals["eq"], [-3.0, -1.0], atol=TOL)
    np.testing.assert_allclose(st.vector_duals["nn"], [0.0, 2.0, 0.0], atol=TOL)
    np.testing.assert_allclose(st.vector_primal, [0.0, 0.0], atol=TOL)
    st = moi_cases.conic_linear_nonpositives(oracle_optimizer())[0]
    np.testing.assert_allclose(st.vector_duals["c"], [-1.0, -1.0], atol=TOL)


def test_moi_certificate_signs():
    """The dual of the constraint that makes each one-variable model
    infeasible has the sign of a Farkas ray in MOI's convention (>= 0 for a
    ">=" side, <= 0 for a "<=" side)."""
    for name, sign in (("dual_certificate_equal_to_lower", -1), ("dual_certificate_equal_to_upper", 1),
                       ("dual_certificate_greater_than", 1), ("dual_certificate_less_than", -1),
                       ("dual_certificate_interval_lower", 1), ("dual_certificate_interval_upper", -1),
                       ("dual_certificate_variable_less_than", 1),
                       ("dual_certificate_variable_less_than_max", 1)):
        st = moi_cases.CASES[name](oracle_optimizer())[0]
        _, d = st.certificate
        assert d * sign > 0, (name, d)


def test_moi_result_count_and_index_bounds():
    opt = oracle_optimizer()
    x = opt.add_variable(0.0, 1.0)
    opt.add_constraint({x: 1.0}, ">=", 0.5)
    opt.set_objective({x: 1.0})
    assert opt.result_count() == 0
    assert opt.termination_status() == "OPTIMIZE_NOT_CALLED"
    assert opt.primal_status() == "NO_SOLUTION" and opt.dual_status() == "NO_SOLUTION"
    assert opt.raw_status_string() == "Optimize not called"
    with pytest.raises(ResultIndexBoundsError):
        opt.objective_value()
    opt.optimize()
    assert opt.result_count() == 1
    assert opt.primal_status(2) == "NO_SOLUTION" and opt.dual_status(2) == "NO_SOLUTION"
    with pytest.raises(ResultIndexBoundsError):
        opt.objective_value(2)
    with pytest.raises(ResultIndexBoundsError):
        opt.variable_primal(x, 2)
    assert opt.raw_status_string() == "TERMINATION_REASON_OPTIMAL : OPTIMAL"
    assert opt.solve_time_sec() >= 0.0
    assert opt.solver_name() == "cuPDLP"
    opt.delete_variable(x)
    with pytest.raises(InvalidIndex):
        opt.set_bounds(x, lower=0.0)
    with pytest.raises(ValueError):
        opt.add_constraint({}, "!=", 0.0)
    with pytest.raises(ValueError):
        opt.add_constraint({}, "interval", (1.0, 0.0))


def test_moi_feasibility_sense():
    st = moi_cases.feasibility_sense(oracle_optimizer())[0]
    assert st.feasible and st.objective == 0.0


def test_moi_silent_keeps_the_parameters():
    opt = oracle_optimizer()
    opt.set_raw("verbosity", 2)
    seen = []
    opt._solve_fn = lambda params, problem: seen.append(params.verbosity) or O.solve(problem, params)
    x = opt.add_variable(0.0, math.inf)
    opt.add_constraint({x: 1.0}, ">=", 1.0)
    opt.set_objective({x: 1.0})
    opt.optimize()
    assert seen == [0] and opt.get_raw("verbosity") == 2
    opt.silent = False
    opt.optimize()
    assert seen == [0, 2]
