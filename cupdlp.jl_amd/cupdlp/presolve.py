"""Presolve of src/preprocess.jl:105-339 (host side; ``optimize`` does not call
it, as in the reference).

``presolve`` removes empty rows (the reference's infeasibility error for an
empty row whose right-hand side cannot hold) and, for a linear objective,
empty columns (each moved to its best bound in the objective constant),
counts singleton rows when verbose, and can turn finite variable bounds into
rows; ``undo_presolve`` maps a solution of the reduced problem back.  Row and
column indices are 0-based here (1-based in the reference).  The reference's
``undo_presolve`` calls a ``projection!`` it never defines
(preprocess.jl:313); here that step is the projection of the recovered primal
point onto the saved variable bounds, which its arguments name.
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .types import QuadraticProgrammingProblem


@dataclass
class PresolveInfo:
    original_primal_size: int
    original_dual_size: int
    empty_rows: np.ndarray
    empty_columns: np.ndarray
    variable_lower_bound: np.ndarray
    variable_upper_bound: np.ndarray


def remove_empty_rows(problem: QuadraticProgrammingProblem) -> np.ndarray:
    """preprocess.jl:112-134."""
    A = problem.constraint_matrix
    seen = np.zeros(A.shape[0], dtype=bool)
    seen[A.indices] = True
    empty = np.flatnonzero(~seen)
    neq = problem.num_equalities
    b = problem.right_hand_side
    for row in empty:
        if row >= neq and b[row] > 0.0:
            raise ValueError("The problem is infeasible.")
        if row < neq and b[row] != 0.0:
            raise ValueError("The problem is infeasible.")
    if empty.size:
        problem.constraint_matrix = sp.csc_matrix(A[seen, :])
        problem.right_hand_side = b[seen]
        problem.num_equalities = neq - int(np.sum(empty < neq))
    return empty


def remove_empty_columns(problem: QuadraticProgrammingProblem) -> np.ndarray:
    """preprocess.jl:145-175 (linear objective)."""
    A = problem.constraint_matrix
    is_empty = np.diff(A.indptr) == 0
    empty = np.flatnonzero(is_empty)
    if empty.size == 0:
        return empty
    c = problem.objective_vector
    for col in empty:
        coef = c[col]
        bound = problem.variable_lower_bound[col] if coef >= 0 else problem.variable_upper_bound[col]
        problem.objective_constant += bound * coef
    keep = ~is_empty
    problem.constraint_matrix = sp.csc_matrix(A[:, keep])
    problem.objective_vector = c[keep]
    problem.variable_lower_bound = problem.variable_lower_bound[keep]
    problem.variable_upper_bound = problem.variable_upper_bound[keep]
    return empty


def transform_bounds_into_linear_constraints(qp: QuadraticProgrammingProblem) -> None:
    """preprocess.jl:180-209: rows x_j >= l_j (finite l), then -x_j >= -u_j
    (finite u), appended; the bounds become infinite."""
    lo = np.flatnonzero(np.isfinite(qp.variable_lower_bound))
    hi = np.flatnonzero(np.isfinite(qp.variable_upper_bound))
    k = lo.size + hi.size
    n = qp.variable_lower_bound.size
    block = sp.csc_matrix((np.concatenate([np.ones(lo.size), -np.ones(hi.size)]),
                           (np.arange(k), np.concatenate([lo, hi]))), shape=(k, n))
    qp.constraint_matrix = sp.csc_matrix(sp.vstack([qp.constraint_matrix, block]))
    qp.right_hand_side = np.concatenate([qp.right_hand_side, qp.variable_lower_bound[lo],
                                         -qp.variable_upper_bound[hi]])
    qp.variable_lower_bound = np.full(n, -np.inf)
    qp.variable_upper_bound = np.full(n, np.inf)


def check_for_singleton_constraints(qp: QuadraticProgrammingProblem) -> int:
    """preprocess.jl:259-273: reports the rows with exactly one entry."""
    per_row = np.bincount(qp.constraint_matrix.indices, minlength=qp.constraint_matrix.shape[0])
    num_single = int(np.sum(per_row == 1))
    if num_single > 0:
        print(f"{num_single} constraints involving exactly a single variable")
    return num_single


def presolve(qp: QuadraticProgrammingProblem, verbosity: int = 1,
             transform_bounds: bool = False) -> PresolveInfo:
    """preprocess.jl:223-257: modifies qp in place."""
    saved_l = qp.variable_lower_bound.copy()
    saved_u = qp.variable_upper_bound.copy()
    original_dual_size, original_primal_size = qp.constraint_matrix.shape
    empty_rows = remove_empty_rows(qp)
    # the GPU path has no quadratic term, so the objective is always linear here
    empty_columns = remove_empty_columns(qp)
    if verbosity >= 1:
        check_for_singleton_constraints(qp)
    if transform_bounds:
        transform_bounds_into_linear_constraints(qp)
    return PresolveInfo(original_primal_size, original_dual_size, empty_rows, empty_columns,
                        saved_l, saved_u)


def recover_original_solution(solution, empty_indices, original_size) -> np.ndarray:
    """preprocess.jl:285-299: zeros at the removed indices."""
    keep = np.ones(original_size, dtype=bool)
    keep[np.asarray(empty_indices, dtype=np.int64)] = False
    out = np.zeros(original_size)
    out[keep] = np.asarray(solution, dtype=float)[:int(keep.sum())]
    return out


def undo_presolve(info: PresolveInfo, primal_solution, dual_solution):
    """preprocess.jl:301-327: the solution of the original problem."""
    x = recover_original_solution(primal_solution, info.empty_columns, info.original_primal_size)
    x = np.minimum(info.variable_upper_bound, np.maximum(info.variable_lower_bound, x))
    y = recover_original_solution(dual_solution, info.empty_rows, info.original_dual_size)
    return x, y
