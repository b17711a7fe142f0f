"""Presolve (src/preprocess.jl:105-339) against hand-derived results, and a
presolve -> solve -> undo round trip on the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from cupdlp import PdhgParameters, QuadraticProgrammingProblem, construct_termination_criteria
from cupdlp.presolve import presolve, transform_bounds_into_linear_constraints, undo_presolve

INF = np.inf


def lp(A, b, c, l, u, neq, c0=0.0):
    return QuadraticProgrammingProblem(np.array(l, float), np.array(u, float), None,
                                       np.array(c, float), c0, sp.csc_matrix(np.array(A, float)),
                                       np.array(b, float), neq)


def example():
    """Rows: r0 (eq, empty, b 0), r1 (eq: x0 + x2 = 2), r2 (>=, empty, b -1),
    r3 (>=: x0 - x2 >= -1, a singleton? no), r4 (>=: x2 >= 0.5, a singleton).
    Columns: x1 empty with c 3 >= 0 (to its lower bound 1: +3), x3 empty with
    c -2 < 0 (to its upper bound 4: -8), x0 and x2 used."""
    A = [[0, 0, 0, 0],
         [1, 0, 1, 0],
         [0, 0, 0, 0],
         [1, 0, -1, 0],
         [0, 0, 1, 0]]
    return lp(A, [0, 2, -1, -1, 0.5], [1, 3, 2, -2], [0, 1, 0, -INF], [10, 5, 10, 4], 2, c0=0.5)


def test_presolve_hand_derived(capsys):
    p = example()
    info = presolve(p)
    assert list(info.empty_rows) == [0, 2] and list(info.empty_columns) == [1, 3]
    assert info.original_dual_size == 5 and info.original_primal_size == 4
    # rows r1 (eq), r3, r4 remain; one equality removed
    assert p.num_equalities == 1
    assert np.array_equal(p.constraint_matrix.toarray(), [[1, 1], [1, -1], [0, 1]])
    assert list(p.right_hand_side) == [2, -1, 0.5]
    assert list(p.objective_vector) == [1, 2]
    assert list(p.variable_lower_bound) == [0, 0] and list(p.variable_upper_bound) == [10, 10]
    assert p.objective_constant == 0.5 + 3.0 - 8.0
    assert "1 constraints involving exactly a single variable" in capsys.readouterr().out
    # undo: zeros at the removed indices, then the saved bounds
    x, y = undo_presolve(info, [1.5, 0.5], [7.0, 8.0, 9.0])
    assert list(x) == [1.5, 1.0, 0.5, 0.0]  # x1 := max(1, 0) = 1; x3 := min(4, 0) = 0
    assert list(y) == [0.0, 7.0, 0.0, 8.0, 9.0]


def test_presolve_infeasible_empty_rows():
    with pytest.raises(ValueError, match="infeasible"):
        presolve(lp([[0.0], [1.0]], [1.0, 0.0], [1.0], [0], [1], 1))   # empty equality, b != 0
    with pytest.raises(ValueError, match="infeasible"):
        presolve(lp([[1.0], [0.0]], [0.0, 0.5], [1.0], [0], [1], 1))   # empty >= row, b > 0


def test_transform_bounds():
    p = lp([[1.0, 1.0, 0.0]], [1.0], [1, 1, 1], [0, -INF, -2], [3, INF, INF], 0)
    transform_bounds_into_linear_constraints(p)
    # lower-bound rows (x0 >= 0, x2 >= -2), then upper-bound rows (-x0 >= -3)
    assert np.array_equal(p.constraint_matrix.toarray(),
                          [[1, 1, 0], [1, 0, 0], [0, 0, 1], [-1, 0, 0]])
    assert list(p.right_hand_side) == [1, 0, -2, -3]
    assert np.all(np.isneginf(p.variable_lower_bound)) and np.all(np.isposinf(p.variable_upper_bound))


def test_presolve_solve_undo_round_trip():
    """The optimum of the original LP equals the reduced one's plus the
    constant, and the recovered point is feasible for the original."""
    orig = example()
    red = example()
    info = presolve(red, verbosity=0)
    tc = construct_termination_criteria(eps_optimal_absolute=1e-9, eps_optimal_relative=1e-9)
    p = PdhgParameters(termination_criteria=tc, record_iteration_stats=False)
    a, r = O.solve(orig, p), O.solve(red, p)
    assert a.termination_string == "OPTIMAL" and r.termination_string == "OPTIMAL"
    obj_orig = float(orig.objective_vector @ a.primal_solution) + orig.objective_constant
    obj_red = float(red.objective_vector @ r.primal_solution) + red.objective_constant
    assert abs(obj_orig - obj_red) <= 1e-6 * (1 + abs(obj_orig))
    x, _ = undo_presolve(info, r.primal_solution, r.dual_solution)
    assert np.all(x >= orig.variable_lower_bound) and np.all(x <= orig.variable_upper_bound)
    A = orig.constraint_matrix
    Ax = A @ x
    assert np.allclose(Ax[:2], orig.right_hand_side[:2], atol=1e-6)
    assert np.all(Ax[2:] >= orig.right_hand_side[2:] - 1e-6)
