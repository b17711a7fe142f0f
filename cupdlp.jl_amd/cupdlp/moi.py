"""Model front end mirroring cuPDLP.Optimizer (src/MOI_wrapper.jl).

The reference is a MathOptInterface optimizer; MOI is Julia-only, so this is a
small model-building API with the same semantics:

* equality rows are stored first, then ">=" rows, each in insertion order
  (the _LPProductOfSets cache, src/MOI_wrapper.jl:4-24);
* "<=" rows are negated into ">=" rows, as MOI's bridges do, and their duals
  are negated back;
* a MAX objective is negated before the solve and flipped back in the
  getters (src/MOI_wrapper.jl:168-170, 188-189, 243-251);
* RawOptimizerAttribute names are PdhgParameters fields (:104-121);
  ``silent`` zeroes verbosity for the solve (:125-132, 194-201);
* termination / primal / dual status maps (:224-300);
* ConstraintDual of a ">=" row reads dual_solution[num_equalities + index] (:311-318).

Two documented differences: a model with no rows (m = 0) solves instead of
failing in the kernel launch (test/MOI_wrapper.jl:33-34), and
``solve_time_sec`` works (src/MOI_wrapper.jl:212-214 reads a missing field).
"""
import math
from dataclasses import replace

import numpy as np
import scipy.sparse as sp

from .types import PdhgParameters, QuadraticProgrammingProblem, TerminationReason as TR

_TERMINATION_STATUS = {
    TR.TERMINATION_REASON_UNSPECIFIED: "OPTIMIZE_NOT_CALLED",
    TR.TERMINATION_REASON_OPTIMAL: "OPTIMAL",
    TR.TERMINATION_REASON_PRIMAL_INFEASIBLE: "INFEASIBLE",
    TR.TERMINATION_REASON_DUAL_INFEASIBLE: "DUAL_INFEASIBLE",
    TR.TERMINATION_REASON_TIME_LIMIT: "TIME_LIMIT",
    TR.TERMINATION_REASON_ITERATION_LIMIT: "ITERATION_LIMIT",
    TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT: "NUMERICAL_ERROR",
    TR.TERMINATION_REASON_NUMERICAL_ERROR: "NUMERICAL_ERROR",
    TR.TERMINATION_REASON_INVALID_PROBLEM: "INVALID_MODEL",
    TR.TERMINATION_REASON_OTHER: "OTHER_ERROR",
}
_UNKNOWN = "UNKNOWN_RESULT_STATUS"
_PRIMAL_STATUS = {r: _UNKNOWN for r in TR}
_PRIMAL_STATUS.update({TR.TERMINATION_REASON_UNSPECIFIED: "NO_SOLUTION",
                       TR.TERMINATION_REASON_OPTIMAL: "FEASIBLE_POINT",
                       TR.TERMINATION_REASON_PRIMAL_INFEASIBLE: "NO_SOLUTION",
                       TR.TERMINATION_REASON_DUAL_INFEASIBLE: "INFEASIBILITY_CERTIFICATE"})
_DUAL_STATUS = {r: _UNKNOWN for r in TR}
_DUAL_STATUS.update({TR.TERMINATION_REASON_UNSPECIFIED: "NO_SOLUTION",
                     TR.TERMINATION_REASON_OPTIMAL: "FEASIBLE_POINT",
                     TR.TERMINATION_REASON_PRIMAL_INFEASIBLE: "INFEASIBILITY_CERTIFICATE",
                     TR.TERMINATION_REASON_DUAL_INFEASIBLE: "NO_SOLUTION"})


class ConstraintIndex:
    __slots__ = ("kind", "value", "negated")

    def __init__(self, kind, value, negated=False):
        self.kind, self.value, self.negated = kind, value, negated

    def __repr__(self):
        return f"ConstraintIndex({self.kind}, {self.value})"


class Optimizer:
    """cuPDLP.Optimizer: defaults of src/MOI_wrapper.jl:52-80."""

    def __init__(self, solve_fn=None, **raw):
        self.parameters = PdhgParameters()
        self.silent = False
        self.output = None
        self.max_sense = False
        self.num_equalities = 0
        self._solve_fn = solve_fn
        self._lb, self._ub = [], []
        self._rows = {"eq": [], "ge": []}  # list of (dict terms, rhs)
        self._obj_terms, self._obj_const = {}, 0.0
        for k, v in raw.items():
            self.set_raw(k, v)

    # -- attributes --
    def set_raw(self, name, value):
        if not hasattr(self.parameters, name):
            raise KeyError(f"UnsupportedAttribute: {name}")
        setattr(self.parameters, name, value)

    def get_raw(self, name):
        if not hasattr(self.parameters, name):
            raise KeyError(f"UnsupportedAttribute: {name}")
        return getattr(self.parameters, name)

    def solver_name(self):
        return "cuPDLP"

    def is_empty(self):
        return self.output is None

    def empty(self):
        self.output = None

    # -- model building --
    def add_variable(self, lower=-math.inf, upper=math.inf):
        self._lb.append(float(lower)); self._ub.append(float(upper))
        return len(self._lb) - 1

    def add_variables(self, k):
        return [self.add_variable() for _ in range(k)]

    def set_bounds(self, vi, lower=None, upper=None):
        if lower is not None:
            self._lb[vi] = float(lower)
        if upper is not None:
            self._ub[vi] = float(upper)

    def add_constraint(self, terms, sense, rhs):
        """terms: {variable index: coefficient}; sense: '==', '>=' or '<='."""
        terms = dict(terms)
        if sense == "==":
            self._rows["eq"].append((terms, float(rhs)))
            return ConstraintIndex("eq", len(self._rows["eq"]) - 1)
        if sense == ">=":
            self._rows["ge"].append((terms, float(rhs)))
            return ConstraintIndex("ge", len(self._rows["ge"]) - 1)
        if sense == "<=":
            self._rows["ge"].append(({k: -v for k, v in terms.items()}, -float(rhs)))
            return ConstraintIndex("ge", len(self._rows["ge"]) - 1, negated=True)
        raise ValueError(f"unsupported constraint sense {sense!r}")

    def set_objective(self, terms, constant=0.0, sense="min"):
        self._obj_terms = dict(terms)
        self._obj_const = float(constant)
        self.max_sense = sense == "max"

    def _flip(self, v):
        return -v if self.max_sense else v

    def to_problem(self) -> QuadraticProgrammingProblem:
        n = len(self._lb)
        rows = self._rows["eq"] + self._rows["ge"]
        r, cidx, val = [], [], []
        for i, (terms, _) in enumerate(rows):
            for j, a in terms.items():
                r.append(i); cidx.append(j); val.append(a)
        A = sp.csc_matrix((val, (r, cidx)), shape=(len(rows), n))
        c = np.zeros(n)
        for j, a in self._obj_terms.items():
            c[j] += a
        self.num_equalities = len(self._rows["eq"])
        return QuadraticProgrammingProblem(
            np.array(self._lb, dtype=float), np.array(self._ub, dtype=float), None,
            self._flip(c), self._flip(self._obj_const), A,
            np.array([rhs for _, rhs in rows], dtype=float), self.num_equalities)

    def optimize(self):
        self.empty()
        problem = self.to_problem()
        params = self.parameters
        if self.silent:
            params = replace(params, verbosity=0)
        if self._solve_fn is not None:
            self.output = self._solve_fn(params, problem)
        else:
            from .solver import optimize
            self.output = optimize(params, problem)
        return self.output

    # -- results --
    def result_count(self):
        return 0 if self.output is None else 1

    def termination_status(self):
        return "OPTIMIZE_NOT_CALLED" if self.output is None else \
            _TERMINATION_STATUS[self.output.termination_reason]

    def primal_status(self):
        return "NO_SOLUTION" if self.output is None else _PRIMAL_STATUS[self.output.termination_reason]

    def dual_status(self):
        return "NO_SOLUTION" if self.output is None else _DUAL_STATUS[self.output.termination_reason]

    def raw_status_string(self):
        if self.output is None:
            return "Optimize not called"
        return f"{self.output.termination_reason.name} : {self.output.termination_string}"

    def objective_value(self):
        return self._flip(self.output.iteration_stats[-1].convergence_information[0].primal_objective)

    def dual_objective_value(self):
        return self._flip(self.output.iteration_stats[-1].convergence_information[0].dual_objective)

    def variable_primal(self, vi):
        return float(self.output.primal_solution[vi])

    def constraint_dual(self, ci: ConstraintIndex):
        if ci.kind == "eq":
            v = self.output.dual_solution[ci.value]
        else:
            v = self.output.dual_solution[self.num_equalities + ci.value]
        return float(-v if ci.negated else v)

    def solve_time_sec(self):
        return self.output.iteration_stats[-1].cumulative_time_sec
