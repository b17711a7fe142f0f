"""Model front end mirroring cuPDLP.Optimizer (src/MOI_wrapper.jl).

The reference is a MathOptInterface optimizer used behind MOI's caching
optimizer and bridges (test/MOI_wrapper.jl:9-16).  MOI is Julia-only, so this
is a small model-building API that gives the same answers as that stack:

* the optimizer itself takes equality rows and ">=" rows, equalities first,
  each in insertion order (the _LPProductOfSets cache, src/MOI_wrapper.jl:4-24),
  and variable bounds (BOUND_SETS, :27-32);
* what the bridges add in front of it is done here: a "<=" row is negated into
  a ">=" row and its dual negated back; an interval row lo <= f <= hi becomes
  f >= lo and -f >= -hi (a side at infinity is dropped) and its dual is the sum
  of the two sides' MOI duals; vector rows (f + constant in Nonnegatives,
  Nonpositives or Zeros) become one scalar row per component; repeated terms
  of a function are summed; modifications (coefficients, right-hand sides,
  functions, deletions) edit the cached model, which is copied to the solver
  on every optimize;
* a MAX objective is negated before the solve and flipped back in the
  objective getters (src/MOI_wrapper.jl:168-170, 188-189, 243-251); duals are
  returned as the solver gives them (:302-318);
* RawOptimizerAttribute names are PdhgParameters fields (:104-121);
  ``silent`` zeroes verbosity for the solve (:125-132, 194-201);
* termination / primal / dual status maps (:224-300) and ResultCount (:320-326);
  a result getter with result_index above ResultCount raises, as
  MOI.check_result_index_bounds does.

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

_SCALAR_SENSES = ("==", ">=", "<=", "interval")
_CONES = {"nonnegatives": ">=", "nonpositives": "<=", "zeros": "=="}


class ResultIndexBoundsError(IndexError):
    """A result getter asked for a result beyond ResultCount."""


class InvalidIndex(KeyError):
    """A deleted (or never created) variable or constraint was used."""


class ConstraintIndex:
    """A scalar row (kind "eq", "ge", "le" or "interval") or a vector row
    (kind "vector": its components are scalar rows)."""
    __slots__ = ("kind", "value")

    def __init__(self, kind, value):
        self.kind, self.value = kind, value

    @property
    def negated(self):
        return self.kind == "le"

    def __repr__(self):
        return f"ConstraintIndex({self.kind}, {self.value})"


def _canonical(terms):
    """{variable: coefficient} from a dict or from (variable, coefficient)
    pairs; repeated variables are summed, as MOI canonicalizes a function."""
    out = {}
    items = terms.items() if isinstance(terms, dict) else terms
    for v, a in items:
        out[v] = out.get(v, 0.0) + float(a)
    return out


class Optimizer:
    """cuPDLP.Optimizer: defaults of src/MOI_wrapper.jl:52-80."""

    def __init__(self, solve_fn=None, **raw):
        self.parameters = PdhgParameters()
        self.silent = False
        self.output = None
        self.max_sense = False
        self.num_equalities = 0
        self._solve_fn = solve_fn
        self._vars = []      # [lower, upper, alive]
        self._cons = []      # dicts: kind, terms, lo, hi, alive (+ comps for vector rows)
        self._obj_terms, self._obj_const = {}, 0.0
        self._sense = "min"
        self._col = {}       # variable -> column of the last solve
        self._rows_of = {}   # scalar constraint -> [(row, sign)] of the last solve
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

    # -- variables --
    def add_variable(self, lower=-math.inf, upper=math.inf):
        self._vars.append([float(lower), float(upper), True])
        return len(self._vars) - 1

    def add_variables(self, k):
        return [self.add_variable() for _ in range(k)]

    def _var(self, vi):
        if not (0 <= vi < len(self._vars)) or not self._vars[vi][2]:
            raise InvalidIndex(f"variable {vi}")
        return self._vars[vi]

    def is_valid_variable(self, vi):
        return 0 <= vi < len(self._vars) and self._vars[vi][2]

    def set_bounds(self, vi, lower=None, upper=None):
        v = self._var(vi)
        if lower is not None:
            v[0] = float(lower)
        if upper is not None:
            v[1] = float(upper)

    def bounds(self, vi):
        v = self._var(vi)
        return v[0], v[1]

    def delete_variable(self, vi):
        self._var(vi)[2] = False
        self._obj_terms.pop(vi, None)
        for c in self._cons:
            for t in c.get("comps", [c]):
                t["terms"].pop(vi, None)

    def delete_variables(self, vis):
        for vi in vis:
            self._var(vi)
        for vi in vis:
            self.delete_variable(vi)

    def num_variables(self):
        return sum(1 for v in self._vars if v[2])

    # -- constraints --
    def _check_terms(self, terms):
        for vi in terms:
            self._var(vi)
        return terms

    def add_constraint(self, terms, sense, rhs):
        """terms: {variable: coefficient} or (variable, coefficient) pairs;
        sense "==", ">=", "<=" (rhs a number) or "interval" (rhs = (lo, hi))."""
        if sense not in _SCALAR_SENSES:
            raise ValueError(f"unsupported constraint sense {sense!r}")
        terms = self._check_terms(_canonical(terms))
        lo, hi = self._set_bounds(sense, rhs)
        kind = {"==": "eq", ">=": "ge", "<=": "le", "interval": "interval"}[sense]
        self._cons.append({"kind": kind, "terms": terms, "lo": lo, "hi": hi, "alive": True})
        return ConstraintIndex(kind, len(self._cons) - 1)

    @staticmethod
    def _set_bounds(sense, rhs):
        if sense == "interval":
            lo, hi = (float(rhs[0]), float(rhs[1]))
            if lo > hi:
                raise ValueError("interval with lower > upper")
            return lo, hi
        r = float(rhs)
        return {"==": (r, r), ">=": (r, math.inf), "<=": (-math.inf, r)}[sense]

    def add_vector_constraint(self, rows, constants, cone):
        """Each component i: rows[i] (terms) + constants[i] in the cone
        ("nonnegatives", "nonpositives" or "zeros")."""
        if cone not in _CONES:
            raise ValueError(f"unsupported cone {cone!r}")
        if len(rows) != len(constants):
            raise ValueError("one constant per component")
        sense = _CONES[cone]
        comps = []
        for terms, k in zip(rows, constants):
            t = self._check_terms(_canonical(terms))
            lo, hi = self._set_bounds(sense, -float(k))
            kind = {"==": "eq", ">=": "ge", "<=": "le"}[sense]
            comps.append({"kind": kind, "terms": t, "lo": lo, "hi": hi, "alive": True})
        self._cons.append({"kind": "vector", "comps": comps, "alive": True})
        return ConstraintIndex("vector", len(self._cons) - 1)

    def add_variables_in_cone(self, vis, cone):
        """VectorOfVariables in a cone: one component row per variable."""
        return self.add_vector_constraint([{v: 1.0} for v in vis], [0.0] * len(vis), cone)

    def _con(self, ci):
        if not (0 <= ci.value < len(self._cons)) or not self._cons[ci.value]["alive"]:
            raise InvalidIndex(repr(ci))
        return self._cons[ci.value]

    def is_valid_constraint(self, ci):
        return 0 <= ci.value < len(self._cons) and self._cons[ci.value]["alive"]

    def delete_constraint(self, ci):
        self._con(ci)["alive"] = False

    def set_constraint_coefficient(self, ci, vi, a):
        c = self._con(ci)
        self._var(vi)
        if c["kind"] == "vector":
            raise ValueError("vector rows: modify a component with set_constraint_function")
        if a == 0.0:
            c["terms"].pop(vi, None)
        else:
            c["terms"][vi] = float(a)

    def set_constraint_function(self, ci, terms):
        c = self._con(ci)
        if c["kind"] == "vector":
            if len(terms) != len(c["comps"]):
                raise ValueError("one function per component")
            for comp, t in zip(c["comps"], terms):
                comp["terms"] = self._check_terms(_canonical(t))
        else:
            c["terms"] = self._check_terms(_canonical(terms))

    def set_constraint_set(self, ci, rhs):
        """A new right-hand side (a number; (lo, hi) for an interval row)."""
        c = self._con(ci)
        if c["kind"] == "vector":
            raise ValueError("vector rows: their constants are part of the function")
        sense = {"eq": "==", "ge": ">=", "le": "<=", "interval": "interval"}[c["kind"]]
        c["lo"], c["hi"] = self._set_bounds(sense, rhs)

    def constraint_function(self, ci):
        c = self._con(ci)
        if c["kind"] == "vector":
            return [dict(t["terms"]) for t in c["comps"]]
        return dict(c["terms"])

    # -- objective --
    def set_objective(self, terms, constant=0.0, sense="min"):
        self._obj_terms = self._check_terms(_canonical(terms))
        self._obj_const = float(constant)
        self.set_objective_sense(sense)

    def set_objective_sense(self, sense):
        """"min", "max" or "feasibility" (which clears the objective)."""
        if sense not in ("min", "max", "feasibility"):
            raise ValueError(f"unsupported objective sense {sense!r}")
        if sense == "feasibility":
            self._obj_terms, self._obj_const = {}, 0.0
        self._sense = sense
        self.max_sense = sense == "max"

    def set_objective_coefficient(self, vi, a):
        self._var(vi)
        if a == 0.0:
            self._obj_terms.pop(vi, None)
        else:
            self._obj_terms[vi] = float(a)

    def set_objective_constant(self, constant):
        self._obj_const = float(constant)

    def _flip(self, v):
        return -v if self.max_sense else v

    # -- the copy to the solver --
    def to_problem(self) -> QuadraticProgrammingProblem:
        self._col = {}
        lb, ub = [], []
        for vi, (lo, hi, alive) in enumerate(self._vars):
            if alive:
                self._col[vi] = len(lb)
                lb.append(lo)
                ub.append(hi)
        n = len(lb)
        eq, ge = [], []       # (constraint key, terms, rhs, sign)
        for k, c in enumerate(self._cons):
            if not c["alive"]:
                continue
            comps = c["comps"] if c["kind"] == "vector" else [c]
            for p, t in enumerate(comps):
                key = (k, p)
                if t["kind"] == "eq":
                    eq.append((key, t["terms"], t["lo"], 1.0))
                    continue
                if math.isfinite(t["lo"]):
                    ge.append((key, t["terms"], t["lo"], 1.0))
                if math.isfinite(t["hi"]):
                    ge.append((key, {v: -a for v, a in t["terms"].items()}, -t["hi"], -1.0))
        rows = eq + ge
        self._rows_of = {}
        r, cidx, val = [], [], []
        for i, (key, terms, _, sign) in enumerate(rows):
            self._rows_of.setdefault(key, []).append((i, sign))
            for v, a in terms.items():
                if a != 0.0:
                    r.append(i)
                    cidx.append(self._col[v])
                    val.append(a)
        A = sp.csc_matrix((val, (r, cidx)), shape=(len(rows), n))
        A.sum_duplicates()
        c = np.zeros(n)
        for v, a in self._obj_terms.items():
            c[self._col[v]] += a
        self.num_equalities = len(eq)
        return QuadraticProgrammingProblem(
            np.array(lb, dtype=float), np.array(ub, dtype=float), None,
            self._flip(c), self._flip(self._obj_const), A,
            np.array([rhs for _, _, rhs, _ in rows], dtype=float), self.num_equalities)

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

    def _check_result(self, result_index):
        if result_index < 1 or result_index > self.result_count():
            raise ResultIndexBoundsError(f"result {result_index} of {self.result_count()}")

    def termination_status(self):
        return "OPTIMIZE_NOT_CALLED" if self.output is None else \
            _TERMINATION_STATUS[self.output.termination_reason]

    def primal_status(self, result_index=1):
        if result_index > self.result_count():
            return "NO_SOLUTION"
        return _PRIMAL_STATUS[self.output.termination_reason]

    def dual_status(self, result_index=1):
        if result_index > self.result_count():
            return "NO_SOLUTION"
        return _DUAL_STATUS[self.output.termination_reason]

    def raw_status_string(self):
        if self.output is None:
            return "Optimize not called"
        return f"{self.output.termination_reason.name} : {self.output.termination_string}"

    def objective_value(self, result_index=1):
        self._check_result(result_index)
        return self._flip(self.output.iteration_stats[-1].convergence_information[0].primal_objective)

    def dual_objective_value(self, result_index=1):
        self._check_result(result_index)
        return self._flip(self.output.iteration_stats[-1].convergence_information[0].dual_objective)

    def variable_primal(self, vi, result_index=1):
        self._check_result(result_index)
        self._var(vi)
        return float(self.output.primal_solution[self._col[vi]])

    def _scalar_dual(self, key):
        y = self.output.dual_solution
        return float(sum(sign * y[i] for i, sign in self._rows_of.get(key, [])))

    def constraint_dual(self, ci: ConstraintIndex, result_index=1):
        """MOI's dual of the constraint: a "<=" row's is minus its ">=" row's,
        an interval row's the sum over its two sides; a vector row gives one
        value per component."""
        self._check_result(result_index)
        c = self._con(ci)
        if c["kind"] == "vector":
            return [self._scalar_dual((ci.value, p)) for p in range(len(c["comps"]))]
        return self._scalar_dual((ci.value, 0))

    def constraint_primal(self, ci: ConstraintIndex, result_index=1):
        """Value of the constraint's function at the returned point (vector
        rows: with their constants)."""
        self._check_result(result_index)
        c = self._con(ci)
        x = self.output.primal_solution

        def f(terms):
            return float(sum(a * x[self._col[v]] for v, a in terms.items()))
        if c["kind"] == "vector":
            return [f(t["terms"]) - (t["lo"] if t["kind"] != "le" else t["hi"])
                    for t in c["comps"]]
        return f(c["terms"])

    def solve_time_sec(self):
        return self.output.iteration_stats[-1].cumulative_time_sec
