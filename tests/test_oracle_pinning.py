"""Pins the oracle functions the end-to-end checks leave open, each against
the mathematics or the reference's stated semantics rather than the oracle's
own formula:

* initial step size eta0 = 1 / max |A~_ij| (elementwise infinity norm of the
  rescaled matrix), src/primal_dual_hybrid_gradient_gpu.jl:569-571;
* initial primal weight omega0 = primal_importance * ||c~||_2 / ||b~||_2,
  with its guards, src/saddle_point_gpu.jl:767-785;
* the constant step policy's sigma_max estimate against a dense SVD:
  step = 0.8 / sigma_hat with 0.8 sigma_max <= sigma_hat <= sigma_max,
  src/primal_dual_hybrid_gradient_gpu.jl:139-172, 572-583;
* l2_norm_rescaling closed forms, src/preprocess.jl:89-103, 344-358;
* the L_INF optimality norm against residuals computed with numpy,
  src/termination.jl:142-170;
* the iteration / KKT-pass / time limits and their precedence,
  src/termination.jl:236-245, and the KKT-pass accounting (0.5 for the
  initial norm, 1 per step attempt, 2 per evaluation),
  src/primal_dual_hybrid_gradient_gpu.jl:382, 569-570, 632;
* the numerical-error termination through the fault-injection flag
  (movement = 0): the solve stops right after the failing attempt and
  returns the current iterate, src/primal_dual_hybrid_gradient_gpu.jl:384-390,
  627-640, 684-686.
"""
import math

import numpy as np
import pytest

import lp_cases
import oracle as O
from cupdlp import (ConstantStepsizeParams, OptimalityNorm, PdhgParameters, RestartChoice,
                    TerminationReason as TR, construct_termination_criteria)
from cupdlp.generators import config_lp, random_lp, transportation_lp


def params(tol=1e-12, iteration_limit=2**31 - 1, norm=OptimalityNorm.L2,
           kkt_limit=math.inf, time_limit=math.inf, **kw):
    tc = construct_termination_criteria(optimality_norm=norm, eps_optimal_absolute=tol,
                                        eps_optimal_relative=tol, iteration_limit=iteration_limit,
                                        kkt_matrix_pass_limit=kkt_limit, time_sec_limit=time_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def small_lps():
    return [("n=2000 LP", config_lp(0)[0]),
            ("30x40 transportation", transportation_lp(30, 40, seed=4)),
            ("scaled Zipf LP n=5000", config_lp(1, scale=0.005)[0]),
            ("n=300 LP", random_lp(n=300, m=150, seed=21)[0])]


def scaled_dense(prob, vals):
    A = prob.constraint_matrix.copy().astype(np.float64)
    A.data = vals.copy()
    return A.toarray()


# ------------------------------------------------------------- eta0, omega0

@pytest.mark.parametrize("l2,pc", [(False, 1.0), (True, None), (True, 1.0)])
def test_initial_step_size_is_inverse_max_abs_entry(l2, pc):
    for name, prob in small_lps():
        p = params(iteration_limit=0, l2_norm_rescaling=l2, pock_chambolle_alpha=pc)
        vals = O.rescale(prob, p)[0]
        out = O.solve(prob, p)
        assert out.iteration_count == 0
        assert out.iteration_stats[0].step_size == 1.0 / np.max(np.abs(vals)), name


def test_initial_primal_weight_is_norm_ratio():
    for name, prob in small_lps():
        p = params(iteration_limit=0)
        _, c, _, _, b, _, _ = O.rescale(prob, p)
        w = np.linalg.norm(c) / np.linalg.norm(b)
        got = O.solve(prob, p).iteration_stats[0].primal_weight
        assert abs(got - w) <= 4e-16 * w, (name, got, w)
        got = O.solve(prob, params(iteration_limit=0, primal_importance=2.5)).iteration_stats[0]
        assert abs(got.primal_weight - 2.5 * w) <= 8e-16 * w, name
        got = O.solve(prob, params(iteration_limit=0, primal_importance=2.5,
                                   scale_invariant_initial_primal_weight=False))
        assert got.iteration_stats[0].primal_weight == 2.5, name
    # ||b~|| = 0: the guard leaves primal_importance
    prob = lp_cases.dual_infeasible()  # b = 0
    got = O.solve(prob, params(iteration_limit=0, primal_importance=3.0))
    assert got.iteration_stats[0].primal_weight == 3.0


# ----------------------------------------------------------- power method

def power_iterations(dim, rel=0.2, pfail=0.001):
    """Iteration count of the power method's stopping rule,
    src/primal_dual_hybrid_gradient_gpu.jl:139-148, 160-167."""
    eps = 1.0 - (1.0 - rel) ** 2
    k = 0
    while True:
        if k >= 2:
            q = min(0.824, 0.354 / math.sqrt(eps * (k - 1))) * math.sqrt(dim) * (1 - eps) ** (k - 0.5)
        else:
            q = 1.0
        if q <= pfail:
            return k
        k += 1


def test_constant_step_sigma_estimate_within_bound_of_svd():
    for name, prob in small_lps():
        p = params(iteration_limit=0, step_size_policy_params=ConstantStepsizeParams())
        vals = O.rescale(prob, p)[0]
        sigma = np.linalg.norm(scaled_dense(prob, vals), 2)
        st = O.solve(prob, p).iteration_stats[0]
        sigma_hat = 0.8 / st.step_size
        assert 0.8 * sigma * (1 - 1e-12) <= sigma_hat <= sigma * (1 + 1e-12), (name, sigma_hat / sigma)
        # cumulative passes = the power iterations + one evaluation (2)
        assert st.cumulative_kkt_matrix_passes == power_iterations(prob.num_variables) + 2.0, name


# ------------------------------------------------------------ L2 rescaling

L2_ONLY = dict(l_inf_ruiz_iterations=0, l2_norm_rescaling=True, pock_chambolle_alpha=None)


def test_l2_rescaling_diagonal_closed_form():
    prob = lp_cases.lp([[4.0, 0.0], [0.0, 9.0]], [1.0, 1.0], [1.0, 1.0], [0, 0], [10, 10], 2)
    vals, c, l, u, b, E, D = O.rescale(prob, PdhgParameters(**L2_ONLY))
    np.testing.assert_array_equal(vals, [1.0, 1.0])
    np.testing.assert_array_equal(E, [2.0, 3.0])
    np.testing.assert_array_equal(D, [2.0, 3.0])
    np.testing.assert_array_equal(c, [0.5, 1.0 / 3.0])
    np.testing.assert_array_equal(b, [0.5, 1.0 / 3.0])
    np.testing.assert_array_equal(u, [20.0, 30.0])


def test_l2_rescaling_two_by_two_by_hand():
    # rows (3, 4) and (0, 5): row norms 5, 5; columns (3, 0) and (4, 5):
    # norms 3 and sqrt(41).  E = sqrt(row norms), D = sqrt(column norms).
    A = [[3.0, 4.0], [0.0, 5.0]]
    prob = lp_cases.lp(A, [1.0, 2.0], [1.0, 1.0], [0, 0], [10, 10], 1)
    vals, c, l, u, b, E, D = O.rescale(prob, PdhgParameters(**L2_ONLY))
    Ew = np.array([math.sqrt(5.0), math.sqrt(5.0)])
    Dw = np.array([math.sqrt(3.0), 41.0 ** 0.25])
    np.testing.assert_allclose(E, Ew, rtol=4e-16)
    np.testing.assert_allclose(D, Dw, rtol=4e-16)
    # CSC order: a11, a12, a22
    want = [3.0 / (Ew[0] * Dw[0]), 4.0 / (Ew[0] * Dw[1]), 5.0 / (Ew[1] * Dw[1])]
    np.testing.assert_allclose(vals, want, rtol=8e-16)


def test_l2_rescaling_is_sqrt_of_row_and_column_norms():
    prob, _ = random_lp(n=70, m=40, nnz_per_col=4, seed=31)
    A = prob.constraint_matrix.toarray()
    A[5, :] = 0.0  # an empty row and an empty column keep scale 1
    A[:, 9] = 0.0
    prob = lp_cases.lp(A, prob.right_hand_side, prob.objective_vector,
                       prob.variable_lower_bound, prob.variable_upper_bound, prob.num_equalities)
    vals, c, l, u, b, E, D = O.rescale(prob, PdhgParameters(**L2_ONLY))
    rn, cn = np.linalg.norm(A, axis=1), np.linalg.norm(A, axis=0)
    rn[rn == 0] = 1.0
    cn[cn == 0] = 1.0
    np.testing.assert_allclose(E, np.sqrt(rn), rtol=1e-15)
    np.testing.assert_allclose(D, np.sqrt(cn), rtol=1e-15)
    assert E[5] == 1.0 and D[9] == 1.0
    np.testing.assert_allclose(scaled_dense(prob, vals), A / np.outer(E, D), rtol=1e-14, atol=0)


def test_rescaling_with_l2_pass_is_a_diagonal_similarity():
    prob, _ = config_lp(1, scale=0.002)
    p = PdhgParameters(l2_norm_rescaling=True)  # Ruiz, then L2, then Pock-Chambolle
    vals, c, l, u, b, E, D = O.rescale(prob, p)
    A = prob.constraint_matrix.toarray()
    np.testing.assert_allclose(scaled_dense(prob, vals), A / np.outer(E, D), rtol=1e-13, atol=0)
    np.testing.assert_allclose(c, prob.objective_vector / D, rtol=1e-15)
    np.testing.assert_allclose(b, prob.right_hand_side / E, rtol=1e-15)
    fin = np.isfinite(prob.variable_upper_bound)
    np.testing.assert_allclose(u[fin], prob.variable_upper_bound[fin] * D[fin], rtol=1e-15)
    # the L2 pass changes the result (it is not skipped)
    vals0 = O.rescale(prob, PdhgParameters())[0]
    assert not np.array_equal(vals0, vals)


# --------------------------------------------------------------- L_INF norm

def numpy_residuals(prob, x, y):
    """l_inf / l2 primal and dual residuals of (x, y) on the original LP
    (src/iteration_stats_utils_gpu.jl:26-246), with numpy."""
    A = prob.constraint_matrix
    neq = prob.num_equalities
    t = prob.right_hand_side - A @ x
    cv = np.concatenate([t[:neq], np.maximum(t[neq:], 0.0)])
    lo, hi = prob.variable_lower_bound, prob.variable_upper_bound
    pv = np.concatenate([cv, np.maximum(lo - x, 0.0), np.maximum(x - hi, 0.0)])
    g = prob.objective_vector - A.T @ y
    lam = np.maximum(g, 0.0) * np.isfinite(lo) + np.minimum(g, 0.0) * np.isfinite(hi)
    dv = np.concatenate([np.maximum(-y[neq:], 0.0), g - lam])
    return (np.max(np.abs(pv)), np.linalg.norm(pv), np.max(np.abs(dv), initial=0.0),
            np.linalg.norm(dv))


def criteria_met(ci, prob, tol, norm):
    b, c = prob.right_hand_side, prob.objective_vector
    if norm == OptimalityNorm.L_INF:
        pe, pb, de, db = ci.l_inf_primal_residual, np.max(np.abs(b)), ci.l_inf_dual_residual, np.max(np.abs(c))
    else:
        pe, pb, de, db = ci.l2_primal_residual, np.linalg.norm(b), ci.l2_dual_residual, np.linalg.norm(c)
    gap = abs(ci.primal_objective - ci.dual_objective)
    absobj = abs(ci.primal_objective) + abs(ci.dual_objective)
    return de < tol + tol * db and pe < tol + tol * pb and gap < tol + tol * absobj


@pytest.mark.parametrize("norm", [OptimalityNorm.L_INF, OptimalityNorm.L2])
def test_optimality_norm_selects_the_residual_norms(norm):
    prob, _ = config_lp(0, scale=0.5)
    tol = 1e-4
    out = O.solve(prob, params(tol, norm=norm))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    recs = out.iteration_stats
    # terminates at the first evaluation whose criteria in that norm hold
    assert criteria_met(recs[-1].convergence_information[0], prob, tol, norm)
    for r in recs[:-1]:
        assert not criteria_met(r.convergence_information[0], prob, tol, norm), r.iteration_number
    # the recorded residuals are the residuals of the returned point
    li_p, l2_p, li_d, l2_d = numpy_residuals(prob, out.primal_solution, out.dual_solution)
    ci = recs[-1].convergence_information[0]
    np.testing.assert_allclose([ci.l_inf_primal_residual, ci.l2_primal_residual,
                                ci.l_inf_dual_residual, ci.l2_dual_residual],
                               [li_p, l2_p, li_d, l2_d], rtol=1e-7)


def test_l_inf_and_l2_runs_stop_at_their_own_points():
    # evaluated every 4 iterations, so the two tests fire at different points
    prob, _ = config_lp(0, scale=0.5)
    a = O.solve(prob, params(1e-4, norm=OptimalityNorm.L_INF, termination_evaluation_frequency=4))
    b = O.solve(prob, params(1e-4, norm=OptimalityNorm.L2, termination_evaluation_frequency=4))
    assert a.termination_reason == b.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert a.iteration_count != b.iteration_count
    for out, norm in ((a, OptimalityNorm.L_INF), (b, OptimalityNorm.L2)):
        assert criteria_met(out.iteration_stats[-1].convergence_information[0], prob, 1e-4, norm)
        for r in out.iteration_stats[:-1]:
            assert not criteria_met(r.convergence_information[0], prob, 1e-4, norm)


# ------------------------------------------------------------------ limits

def test_kkt_pass_accounting_and_limit():
    prob, _ = config_lp(0, scale=0.25)
    base = O.solve(prob, params(iteration_limit=40))
    recs = base.iteration_stats
    # 0.5 for the initial norm, then 2 per evaluation and 1 per attempt
    assert recs[0].cumulative_kkt_matrix_passes == 2.5
    for r0, r1 in zip(recs, recs[1:]):
        att = r1.method_specific_stats["total_attempts"] - r0.method_specific_stats["total_attempts"]
        assert r1.cumulative_kkt_matrix_passes - r0.cumulative_kkt_matrix_passes == att + 2.0
    # the limit stops at the first evaluation whose passes reach it
    for k in (0, 3, 7):
        lim = recs[k].cumulative_kkt_matrix_passes
        for L in (lim, lim - 0.25):
            out = O.solve(prob, params(iteration_limit=40, kkt_limit=L))
            first = next(r for r in recs if r.cumulative_kkt_matrix_passes >= L)
            assert out.termination_reason == TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT
            assert out.iteration_count == first.iteration_number
            assert out.iteration_stats[-1].cumulative_kkt_matrix_passes == first.cumulative_kkt_matrix_passes


def test_limit_precedence():
    prob, _ = config_lp(0, scale=0.25)
    # time limit 0 holds at the first evaluation
    out = O.solve(prob, params(time_limit=0.0))
    assert out.termination_reason == TR.TERMINATION_REASON_TIME_LIMIT and out.iteration_count == 0
    # iteration limit before KKT-pass limit before time limit (termination.jl:236-245)
    out = O.solve(prob, params(iteration_limit=0, kkt_limit=1.0, time_limit=0.0))
    assert out.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    out = O.solve(prob, params(kkt_limit=1.0, time_limit=0.0))
    assert out.termination_reason == TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT
    # optimality before the limits
    prob, _ = lp_cases.tiny_optimal()
    ref = O.solve(prob, params(1e-6))
    assert ref.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    out = O.solve(prob, params(1e-6, iteration_limit=ref.iteration_count))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL


# --------------------------------------------------------- numerical error

@pytest.mark.parametrize("attempt", [1, 5, 23])
def test_numerical_error_stops_after_the_failing_attempt(attempt):
    prob, _ = config_lp(0, scale=0.25)
    out = O.solve(prob, params(debug_numerical_error_attempt=attempt))
    assert out.termination_reason == TR.TERMINATION_REASON_NUMERICAL_ERROR
    assert out.total_attempts == attempt
    # the take_step! call holding the failing attempt counts as an iteration
    # (it ends by `break`, the loop's counter advances, pdhg:384-390, 626):
    # K - 1 steps were accepted before it
    K = out.iteration_count
    # the returned point is the current iterate (the average is replaced by
    # the current point, pdhg:636-640) ...
    cur = O.solve(prob, params(debug_numerical_error_attempt=attempt), want_current=True)
    np.testing.assert_array_equal(cur.primal_solution, cur.primal_current)
    np.testing.assert_array_equal(cur.dual_solution, cur.dual_current)
    # ... as it stood when the failing step began: after K - 1 accepted steps
    # and the restart decision of the evaluation before step K, which moves
    # the current point to the average on RESTART_TO_AVERAGE only.  A solve
    # stopped at iteration K - 1 holds both candidates.
    ref = O.solve(prob, params(iteration_limit=K - 1), want_current=True)
    assert ref.total_attempts < attempt
    # no record at K - 1: not an evaluation point, so no restart there
    rec = [r for r in out.iteration_stats if r.iteration_number == K - 1]
    choice = rec[0].restart_used if rec else RestartChoice.RESTART_CHOICE_NO_RESTART
    if choice == RestartChoice.RESTART_CHOICE_RESTART_TO_AVERAGE:
        want_x, want_y = ref.primal_solution, ref.dual_solution
    else:
        want_x, want_y = ref.primal_current, ref.dual_current
    np.testing.assert_array_equal(out.primal_solution, want_x)
    np.testing.assert_array_equal(out.dual_solution, want_y)
    # the step size is that of the failing attempt, not updated after it
    last = out.iteration_stats[-1]
    assert last.step_size > 0
    if attempt == 1:
        assert K == 1 and last.step_size == out.iteration_stats[0].step_size


def test_host_relative_kkt_matches_the_oracle_statistics():
    """cupdlp.kkt (numpy, reporting only) against the oracle's statistics of
    the same points: certified optima (0) and returned points."""
    from cupdlp.kkt import kkt_parts
    for prob, cert in (config_lp(0), random_lp(n=500, m=250, seed=3)):
        rp, rd, rg, p, d = kkt_parts(prob, cert.x, cert.y)
        assert max(rp, rd, rg) < 1e-12
        out = O.solve(prob, params(1e-5))
        ci, _ = O.point_stats(prob, out.primal_solution, out.dual_solution, 1e-5, 1e-5)
        rp, rd, rg, p, d = kkt_parts(prob, out.primal_solution, out.dual_solution)
        np.testing.assert_allclose([rp, rd, rg, p, d],
                                   [ci.relative_l2_primal_residual, ci.relative_l2_dual_residual,
                                    ci.relative_optimality_gap, ci.primal_objective,
                                    ci.dual_objective], rtol=1e-8)
