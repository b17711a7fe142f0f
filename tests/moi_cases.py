"""Known-answer models for the MOI-style front end (cupdlp.Optimizer).

Each case follows one of the MathOptInterface.Test functions the reference
runs against cuPDLP.Optimizer (test/tmp.jl:28-67, test/MOI_wrapper.jl:18-65)
and exercises the same feature -- interval rows, vector rows, repeated terms,
modifications and deletions, infeasible and unbounded models with and without
an objective offset, MAX sense, result-count edge cases.  The models are this
project's own small LPs; MOI's test suite itself is Julia and not available
here.  A case is a function that builds a model in an Optimizer, solves it
(possibly several times, editing the model in between) and returns a list of
stages; each stage records the statuses, the expected answer and the solved
problem, so the tests can pin it three ways: the hand-derived answer, HiGHS
on the very problem that was solved, and (on the GPU) the oracle bitwise.
"""
import math

INF = math.inf


class Stage:
    """One solve of a case: what the front end returned, what was expected,
    and the problem that was handed to the solver."""

    def __init__(self, opt, status, objective, primal, duals):
        self.problem = opt.to_problem()
        self.max_sense = opt.max_sense
        self.output = opt.output
        self.termination = opt.termination_status()
        self.primal_status = opt.primal_status()
        self.dual_status = opt.dual_status()
        self.expected_status = status
        self.expected_objective = objective
        self.expected_primal = dict(primal or {})
        self.expected_duals = {k: d for k, (_, d) in (duals or {}).items()}
        ok = opt.result_count() > 0
        self.objective = opt.objective_value() if ok else None
        self.dual_objective = opt.dual_objective_value() if ok else None
        self.primal = {v: opt.variable_primal(v) for v in self.expected_primal}
        self.duals = {k: opt.constraint_dual(ci) for k, (ci, _) in (duals or {}).items()}


def solve(opt, status, objective=None, primal=None, duals=None):
    """Solves and records a stage; duals: {name: (constraint, expected dual)}."""
    opt.optimize()
    return Stage(opt, status, objective, primal, duals)


def linear_integration(opt):
    """test_linear_integration: a model solved, edited and solved again."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 1.0, y: 1.0}, "<=", 1.0)
    opt.set_objective({x: -1.0})
    out = [solve(opt, "OPTIMAL", -1.0, {x: 1.0, y: 0.0}, {"c": (c, -1.0)})]
    opt.set_objective({x: 2.0, y: 1.0}, sense="max")
    out.append(solve(opt, "OPTIMAL", 2.0, {x: 1.0, y: 0.0}))
    z = opt.add_variable(0.0, INF)
    opt.set_constraint_coefficient(c, z, 1.0)
    opt.set_objective_coefficient(z, 3.0)
    out.append(solve(opt, "OPTIMAL", 3.0, {x: 0.0, y: 0.0, z: 1.0}))
    opt.set_constraint_set(c, 2.0)
    out.append(solve(opt, "OPTIMAL", 6.0, {z: 2.0}))
    opt.delete_variable(z)
    out.append(solve(opt, "OPTIMAL", 4.0, {x: 2.0, y: 0.0}))
    opt.set_bounds(x, upper=1.0)
    out.append(solve(opt, "OPTIMAL", 3.0, {x: 1.0, y: 1.0}))
    e = opt.add_constraint({x: 1.0, y: -1.0}, "==", 0.0)
    out.append(solve(opt, "OPTIMAL", 3.0, {x: 1.0, y: 1.0}))
    assert opt.num_equalities == 1 and e.kind == "eq"
    return out


def linear_integration_2(opt):
    """test_linear_integration_2: free variables, an equality, a ">=" row."""
    x = opt.add_variable()
    y = opt.add_variable()
    c1 = opt.add_constraint({x: 1.0, y: 1.0}, "==", 4.0)
    c2 = opt.add_constraint({x: 1.0, y: -1.0}, ">=", 2.0)
    opt.set_objective({x: 1.0, y: 2.0})
    # x = 4 - y and x - y >= 2 give y <= 1; x + 2y = 4 + y is least at y's bound
    opt.set_bounds(y, lower=-1.0)
    return [solve(opt, "OPTIMAL", 3.0, {x: 5.0, y: -1.0}, {"c1": (c1, 1.0), "c2": (c2, 0.0)})]


def linear_integration_delete_variables(opt):
    """test_linear_integration_delete_variables: deleting a batch of variables."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    z = opt.add_variable(0.0, INF)
    opt.add_constraint({x: 1.0, y: 1.0, z: 1.0}, ">=", 1.0)
    opt.set_objective({x: 1.0, y: 2.0, z: 3.0})
    out = [solve(opt, "OPTIMAL", 1.0, {x: 1.0, y: 0.0, z: 0.0})]
    opt.delete_variables([x, y])
    out.append(solve(opt, "OPTIMAL", 3.0, {z: 1.0}))
    assert not opt.is_valid_variable(x) and opt.num_variables() == 1
    return out


def linear_interval_inactive(opt):
    """test_linear_Interval_inactive: an interval row that does not bind."""
    x = opt.add_variable(1.0, INF)
    y = opt.add_variable(2.0, INF)
    c = opt.add_constraint({x: 1.0, y: 1.0}, "interval", (0.0, 10.0))
    opt.set_objective({x: 1.0, y: 1.0})
    return [solve(opt, "OPTIMAL", 3.0, {x: 1.0, y: 2.0}, {"c": (c, 0.0)})]


def linear_interval_active(opt):
    """An interval row binding at its lower, then (MAX) at its upper side."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 1.0, y: 2.0}, "interval", (2.0, 6.0))
    opt.set_objective({x: 2.0, y: 1.0})
    out = [solve(opt, "OPTIMAL", 1.0, {x: 0.0, y: 1.0}, {"c": (c, 0.5)})]
    opt.set_objective({x: 1.0, y: 1.0}, sense="max")
    # max x + y with x + 2y <= 6: x = 6, y = 0; the solver's dual is that of
    # the negated (min) problem, returned as it is: -1 on the upper side
    out.append(solve(opt, "OPTIMAL", 6.0, {x: 6.0, y: 0.0}, {"c": (c, -1.0)}))
    return out


def linear_variable_open_intervals(opt):
    """test_linear_variable_open_intervals: one-sided variable bounds."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(-INF, 0.0)
    opt.add_constraint({x: 1.0, y: 1.0}, "==", 1.0)
    opt.set_objective({x: 1.0, y: -1.0})
    return [solve(opt, "OPTIMAL", 1.0, {x: 1.0, y: 0.0})]


def constraint_duplicate_terms(opt):
    """test_constraint_ScalarAffineFunction_duplicate: x + x >= 1."""
    x = opt.add_variable(0.0, INF)
    c = opt.add_constraint([(x, 1.0), (x, 1.0)], ">=", 1.0)
    opt.set_objective({x: 1.0})
    assert opt.constraint_function(c) == {x: 2.0}
    return [solve(opt, "OPTIMAL", 0.5, {x: 0.5}, {"c": (c, 0.5)})]


def constraint_lessthan(opt):
    """test_constraint_ScalarAffineFunction_LessThan: max x, 2x <= 1."""
    x = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 2.0}, "<=", 1.0)
    opt.set_objective({x: 1.0}, sense="max")
    # the negated problem's ">=" row -2x >= -1 has dual 0.5; a "<=" row's is minus that
    return [solve(opt, "OPTIMAL", 0.5, {x: 0.5}, {"c": (c, -0.5)})]


def conic_linear_vector_affine(opt):
    """test_conic_linear_VectorAffineFunction: rows in Zeros, variables in
    Nonnegatives, as vector-affine functions."""
    x, y, z = opt.add_variables(3)
    eq = opt.add_vector_constraint([{x: 1.0, y: 1.0, z: 1.0}, {y: 1.0, z: 1.0}], [-3.0, -2.0], "zeros")
    nn = opt.add_vector_constraint([{x: 1.0}, {y: 1.0}, {z: 1.0}], [0.0, 0.0, 0.0], "nonnegatives")
    opt.set_objective({x: -3.0, y: -2.0, z: -4.0})
    st = solve(opt, "OPTIMAL", -11.0, {x: 1.0, y: 0.0, z: 2.0})
    st.vector_duals = {"eq": opt.constraint_dual(eq), "nn": opt.constraint_dual(nn)}
    st.vector_primal = opt.constraint_primal(eq)
    return [st]


def conic_linear_vector_of_variables(opt):
    """test_conic_linear_VectorOfVariables: the same with the variables'
    cone given as a vector of variables."""
    x, y, z = opt.add_variables(3)
    opt.add_vector_constraint([{x: 1.0, y: 1.0, z: 1.0}, {y: 1.0, z: 1.0}], [-3.0, -2.0], "zeros")
    opt.add_variables_in_cone([x, y, z], "nonnegatives")
    opt.set_objective({x: -3.0, y: -2.0, z: -4.0})
    return [solve(opt, "OPTIMAL", -11.0, {x: 1.0, y: 0.0, z: 2.0})]


def conic_linear_nonpositives(opt):
    """test_conic_linear_VectorAffineFunction_2 style: a Nonpositives row."""
    x, y = opt.add_variables(2)
    c = opt.add_vector_constraint([{x: -1.0}, {y: -1.0}], [-2.0, -1.0], "nonpositives")
    opt.set_objective({x: 1.0, y: 1.0})
    st = solve(opt, "OPTIMAL", -3.0, {x: -2.0, y: -1.0})
    st.vector_duals = {"c": opt.constraint_dual(c)}
    return [st]


def _infeasible(opt, sense, offset, affine):
    x = opt.add_variable(0.0, INF)
    if affine:  # both conditions are rows
        opt.add_constraint({x: 1.0}, ">=", 2.0)
        opt.add_constraint({x: 2.0}, "<=", 2.0)
    else:       # a bound and a row
        opt.set_bounds(x, lower=2.0)
        opt.add_constraint({x: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0}, offset, sense=sense)
    return [solve(opt, "INFEASIBLE")]


def infeasible_min(opt):
    """test_infeasible_MIN_SENSE."""
    return _infeasible(opt, "min", 0.0, False)


def infeasible_max_offset(opt):
    """test_infeasible_MAX_SENSE_offset."""
    return _infeasible(opt, "max", 1.5, False)


def infeasible_affine_min_offset(opt):
    """test_infeasible_affine_MIN_SENSE_offset."""
    return _infeasible(opt, "min", -2.5, True)


def infeasible_affine_max(opt):
    """test_infeasible_affine_MAX_SENSE."""
    return _infeasible(opt, "max", 0.0, True)


def linear_infeasible(opt):
    """test_linear_INFEASIBLE: x, y >= 0 and x + y <= -1."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    opt.add_constraint({x: 1.0, y: 1.0}, "<=", -1.0)
    opt.set_objective({x: 1.0, y: 1.0})
    return [solve(opt, "INFEASIBLE")]


def linear_infeasible_2(opt):
    """test_linear_INFEASIBLE_2: two rows that contradict."""
    x = opt.add_variable()
    y = opt.add_variable()
    opt.add_constraint({x: 1.0, y: 1.0}, ">=", 2.0)
    opt.add_constraint({x: 1.0, y: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0, y: 1.0})
    return [solve(opt, "INFEASIBLE")]


def _unbounded(opt, sense, offset):
    x = opt.add_variable(0.0, INF)
    opt.add_constraint({x: 1.0}, ">=", 0.0)
    opt.set_objective({x: -1.0 if sense == "min" else 1.0}, offset, sense=sense)
    return [solve(opt, "DUAL_INFEASIBLE")]


def unbounded_min_offset(opt):
    """test_unbounded_MIN_SENSE_offset."""
    return _unbounded(opt, "min", 3.0)


def unbounded_max_offset(opt):
    """test_unbounded_MAX_SENSE_offset."""
    return _unbounded(opt, "max", -3.0)


def _certificate(opt, bound, sense, rhs, max_sense=False):
    x = opt.add_variable(*bound)
    c = opt.add_constraint({x: 1.0}, sense, rhs)
    opt.set_objective({x: 1.0}, sense="max" if max_sense else "min")
    st = solve(opt, "INFEASIBLE")
    st.certificate = (c, opt.constraint_dual(c))
    return [st]


def dual_certificate_equal_to_lower(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_EqualTo_lower."""
    return _certificate(opt, (0.0, INF), "==", -1.0)


def dual_certificate_equal_to_upper(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_EqualTo_upper."""
    return _certificate(opt, (-INF, 0.0), "==", 1.0)


def dual_certificate_greater_than(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_GreaterThan."""
    return _certificate(opt, (-INF, 0.0), ">=", 1.0)


def dual_certificate_less_than(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_LessThan."""
    return _certificate(opt, (0.0, INF), "<=", -1.0)


def dual_certificate_interval_lower(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_Interval_lower."""
    return _certificate(opt, (-INF, 0.0), "interval", (1.0, 2.0))


def dual_certificate_interval_upper(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_Interval_upper."""
    return _certificate(opt, (0.0, INF), "interval", (-2.0, -1.0))


def dual_certificate_variable_less_than(opt):
    """test_solve_DualStatus_INFEASIBILITY_CERTIFICATE_VariableIndex_LessThan."""
    return _certificate(opt, (-INF, -1.0), ">=", 0.0)


def dual_certificate_variable_less_than_max(opt):
    """..._VariableIndex_LessThan_max."""
    return _certificate(opt, (-INF, -1.0), ">=", 0.0, max_sense=True)


def modification_coef_scalar_objective(opt):
    """test_modification_coef_scalar_objective."""
    x = opt.add_variable(1.0, INF)
    opt.add_constraint({x: 1.0}, ">=", 0.0)
    opt.set_objective({x: 1.0})
    out = [solve(opt, "OPTIMAL", 1.0, {x: 1.0})]
    opt.set_objective_coefficient(x, 3.0)
    out.append(solve(opt, "OPTIMAL", 3.0, {x: 1.0}))
    return out


def modification_coef_scalaraffine_lessthan(opt):
    """test_modification_coef_scalaraffine_lessthan."""
    x = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0}, sense="max")
    out = [solve(opt, "OPTIMAL", 1.0, {x: 1.0})]
    opt.set_constraint_coefficient(c, x, 2.0)
    out.append(solve(opt, "OPTIMAL", 0.5, {x: 0.5}))
    return out


def modification_func_scalaraffine_lessthan(opt):
    """test_modification_func_scalaraffine_lessthan."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 1.0}, "<=", 1.0)
    opt.add_constraint({y: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0, y: 1.0}, sense="max")
    out = [solve(opt, "OPTIMAL", 2.0, {x: 1.0, y: 1.0})]
    opt.set_constraint_function(c, {x: 2.0, y: 1.0})
    out.append(solve(opt, "OPTIMAL", 1.0, {x: 0.0, y: 1.0}))
    return out


def modification_set_scalaraffine_lessthan(opt):
    """test_modification_set_scalaraffine_lessthan."""
    x = opt.add_variable(0.0, INF)
    c = opt.add_constraint({x: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0}, sense="max")
    out = [solve(opt, "OPTIMAL", 1.0, {x: 1.0})]
    opt.set_constraint_set(c, 2.0)
    out.append(solve(opt, "OPTIMAL", 2.0, {x: 2.0}))
    return out


def modification_delete_variable_with_objective(opt):
    """test_modification_delete_variable_with_single_variable_obj."""
    x = opt.add_variable(0.0, INF)
    y = opt.add_variable(0.0, INF)
    opt.add_constraint({x: 1.0, y: 1.0}, ">=", 1.0)
    opt.set_objective({x: 2.0, y: 3.0})
    out = [solve(opt, "OPTIMAL", 2.0, {x: 1.0, y: 0.0})]
    opt.delete_variable(x)
    out.append(solve(opt, "OPTIMAL", 3.0, {y: 1.0}))
    return out


def modification_delete_constraint(opt):
    """A deleted row no longer binds."""
    x = opt.add_variable(0.0, 4.0)
    c = opt.add_constraint({x: 1.0}, "<=", 1.0)
    opt.set_objective({x: 1.0}, sense="max")
    out = [solve(opt, "OPTIMAL", 1.0, {x: 1.0})]
    opt.add_constraint({x: 1.0}, ">=", 0.0)
    opt.delete_constraint(c)
    out.append(solve(opt, "OPTIMAL", 4.0, {x: 4.0}))
    assert not opt.is_valid_constraint(c)
    return out


def feasibility_sense(opt):
    """FEASIBILITY_SENSE clears the objective: any feasible point is optimal."""
    x = opt.add_variable(0.0, 1.0)
    opt.add_constraint({x: 1.0}, ">=", 0.25)
    opt.set_objective({x: 5.0}, 2.0)
    opt.set_objective_sense("feasibility")
    st = solve(opt, "OPTIMAL", 0.0)
    st.feasible = (0.25 - 1e-6 <= opt.variable_primal(x) <= 1.0 + 1e-6)
    return [st]


CASES = {f.__name__: f for f in (
    linear_integration, linear_integration_2, linear_integration_delete_variables,
    linear_interval_inactive, linear_interval_active, linear_variable_open_intervals,
    constraint_duplicate_terms, constraint_lessthan,
    conic_linear_vector_affine, conic_linear_vector_of_variables, conic_linear_nonpositives,
    infeasible_min, infeasible_max_offset, infeasible_affine_min_offset, infeasible_affine_max,
    linear_infeasible, linear_infeasible_2, unbounded_min_offset, unbounded_max_offset,
    dual_certificate_equal_to_lower, dual_certificate_equal_to_upper,
    dual_certificate_greater_than, dual_certificate_less_than,
    dual_certificate_interval_lower, dual_certificate_interval_upper,
    dual_certificate_variable_less_than, dual_certificate_variable_less_than_max,
    modification_coef_scalar_objective, modification_coef_scalaraffine_lessthan,
    modification_func_scalaraffine_lessthan, modification_set_scalaraffine_lessthan,
    modification_delete_variable_with_objective, modification_delete_constraint,
    feasibility_sense)}
