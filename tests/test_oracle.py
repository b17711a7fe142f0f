"""Pins the CPU oracle (oracle/pdlp_oracle.c) before it is trusted as the
checker of the GPU path.  Each check is independent of the oracle's own
formulas: closed-form LPs, the generator's optimality certificates, HiGHS
objectives (tests/golden/highs_objectives.json), dense linear algebra,
the Pock-Chambolle operator-norm bound, and hand-derived step-size values.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

import lp_cases
import oracle as O
from cupdlp import (PdhgParameters, RestartChoice, TerminationReason as TR,
                   construct_restart_parameters, construct_termination_criteria)
from cupdlp.generators import config_lp, random_lp, transportation_lp

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "highs_objectives.json")) as f:
    HIGHS = json.load(f)["objectives"]


def params(tol=1e-8, iteration_limit=2**31 - 1, **kw):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def obj(out):
    return out.iteration_stats[-1].convergence_information[0].primal_objective


# ---------------------------------------------------------------- end to end

def test_closed_form_optimal():
    prob, opt = lp_cases.tiny_optimal()
    out = O.solve(prob, params())
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert abs(out.primal_solution[0] - 1.0) < 1e-7
    assert abs(out.dual_solution[0] - 1.0) < 1e-7  # y* = 1 for the row -x >= -1
    assert abs(obj(out) - opt) < 1e-7


def test_closed_form_two_variables():
    prob, opt = lp_cases.tiny_two_var()
    out = O.solve(prob, params())
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    np.testing.assert_allclose(out.primal_solution, [2.0, 0.0], atol=1e-6)
    assert abs(obj(out) - opt) < 1e-6


def test_primal_infeasible_detected_after_iteration_10():
    out = O.solve(lp_cases.primal_infeasible(), params(1e-6))
    assert out.termination_reason == TR.TERMINATION_REASON_PRIMAL_INFEASIBLE
    assert out.iteration_count >= 9  # suppressed while iteration < 10 (pdhg:688-690)


def test_dual_infeasible_detected():
    out = O.solve(lp_cases.dual_infeasible(), params(1e-6))
    assert out.termination_reason == TR.TERMINATION_REASON_DUAL_INFEASIBLE
    assert out.iteration_count >= 9


def test_no_rows_defined():
    prob, opt = lp_cases.no_rows()
    out = O.solve(prob, params())
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    np.testing.assert_allclose(out.primal_solution, [0.0, 2.0], atol=1e-9)


def test_infeasible_and_unbounded_random_lps_detected():
    """x_a + x_b >= 25 with both in [0, 10] makes the LP primal infeasible; a
    free column with cost -1 in no row makes it unbounded (dual infeasible)."""
    out = O.solve(lp_cases.infeasible_random_lp(), params(1e-6))
    assert out.termination_reason == TR.TERMINATION_REASON_PRIMAL_INFEASIBLE
    out = O.solve(lp_cases.unbounded_random_lp(), params(1e-6))
    assert out.termination_reason == TR.TERMINATION_REASON_DUAL_INFEASIBLE


def test_empty_row_and_column_reach_the_certificate():
    """Empty equality and ">=" rows and an empty column, straight to the
    solver: the free dual of the empty equality row and the dual of the
    slack ">=" row stay 0, the empty column goes to its lower bound, and the
    objective is the certificate's."""
    prob, obj = lp_cases.empty_row_and_column_lp()
    out = O.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert out.primal_solution[-1] == 0.0
    assert out.dual_solution[0] == 0.0 and out.dual_solution[-1] == 0.0
    po = out.iteration_stats[-1].convergence_information[0].primal_objective
    assert abs(po - obj) <= 1e-6 * (1 + abs(obj))


def test_inequality_only_and_fixed_variables_reach_the_certificate():
    """No equality rows at all; a tenth of the variables fixed (lb = ub) at
    their certified values; one-sided bounds only: each keeps the
    certificate optimal."""
    prob, cert = random_lp(n=300, m=150, num_eq=0, seed=41)
    out = O.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    po = out.iteration_stats[-1].convergence_information[0].primal_objective
    assert abs(po - cert.objective) <= 1e-6 * (1 + abs(cert.objective))
    prob, cert = lp_cases.fixed_variables_lp()
    out = O.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    np.testing.assert_allclose(out.primal_solution[::10], cert.x[::10], rtol=1e-12, atol=1e-300)
    po = out.iteration_stats[-1].convergence_information[0].primal_objective
    assert abs(po - cert.objective) <= 1e-6 * (1 + abs(cert.objective))
    # one-sided bounds [0, inf) and (-inf, 10] (and free), certificate kept
    prob, cert = lp_cases.one_sided_bounds_lp()
    lb, ub = prob.variable_lower_bound, prob.variable_upper_bound
    assert np.any(np.isfinite(lb) & ~np.isfinite(ub)) and np.any(~np.isfinite(lb) & np.isfinite(ub))
    out = O.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    po = out.iteration_stats[-1].convergence_information[0].primal_objective
    assert abs(po - cert.objective) <= 1e-6 * (1 + abs(cert.objective))


@pytest.mark.parametrize("name,make,tol", [
    ("c0_seed0", lambda: config_lp(0), 1e-8),
    ("random_n300_s11", lambda: random_lp(n=300, m=150, seed=11), 1e-8),
    ("random_n1200_s12", lambda: random_lp(n=1200, m=600, seed=12), 1e-8),
    ("long_row_6000", lambda: lp_cases.long_row_lp(6000), 1e-8),
])
def test_generator_certificate_and_highs(name, make, tol):
    prob, cert_obj = make()
    cert = cert_obj if isinstance(cert_obj, float) else cert_obj.objective
    out = O.solve(prob, params(tol))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    h = HIGHS[name]["objective"]
    assert abs(cert - h) <= 1e-9 * (1 + abs(h))      # certificate is optimal
    assert abs(obj(out) - h) <= 20 * tol * (1 + abs(h))
    # the returned point satisfies the reference optimality test in the original space
    ci, _ = O.point_stats(prob, out.primal_solution, out.dual_solution, tol, tol)
    assert ci.relative_l2_primal_residual < 5 * tol
    assert ci.relative_l2_dual_residual < 5 * tol


def test_zipf_generator_certificate_is_optimal():
    """The Zipf-degree generator of the n = 1,000,000 benchmark LP, at
    n = 20,000: its certificate's objective is HiGHS's optimum, so the many
    PDHG iterations that LP needs (about 474k to 1e-4 here, in the oracle)
    are the algorithm's, not a generator defect."""
    prob, cert = config_lp(1, scale=0.02)
    h = HIGHS["zipf_scaled_n20000"]["objective"]
    assert prob.num_variables == HIGHS["zipf_scaled_n20000"]["n"]
    assert prob.constraint_matrix.nnz == HIGHS["zipf_scaled_n20000"]["nnz"]
    assert abs(cert.objective - h) <= 1e-9 * (1 + abs(h))
    # the certificate point itself is feasible and complementary in FP64
    ci, _ = O.point_stats(prob, cert.x, cert.y, 1e-9, 1e-9)
    assert ci.relative_l2_primal_residual < 1e-9
    assert ci.relative_l2_dual_residual < 1e-9
    assert ci.relative_optimality_gap < 1e-9


def test_transportation_against_highs():
    prob = transportation_lp(60, 60, seed=5)
    out = O.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    h = HIGHS["transport_60x60_s5"]["objective"]
    assert abs(obj(out) - h) <= 1e-6 * abs(h)


def test_iteration_limit_semantics():
    prob, _ = config_lp(0, scale=0.25)
    out = O.solve(prob, params(1e-12, iteration_limit=37))
    assert out.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    assert out.iteration_count == 37
    # evaluations at iterations 1..10 and every 64 (pdhg:627-630), numbered iteration-1
    nums = [s.iteration_number for s in out.iteration_stats]
    assert nums[:10] == list(range(10)) and nums[-1] == 37


def test_restart_schedule_early_artificial_restarts():
    # count >= 0.36 * (iteration - 1) fires at iterations 2 and 3 (count = 1)
    prob, _ = config_lp(0)
    out = O.solve(prob, params(1e-12, iteration_limit=12))
    used = [s.restart_used for s in out.iteration_stats]
    assert used[0] == RestartChoice.RESTART_CHOICE_NO_RESTART   # iteration 1, count 0
    assert used[1] != RestartChoice.RESTART_CHOICE_NO_RESTART
    assert used[2] != RestartChoice.RESTART_CHOICE_NO_RESTART
    assert used[3] == RestartChoice.RESTART_CHOICE_NO_RESTART   # 1 < 0.36 * 3


def test_restart_decision_branches():
    """run_restart_scheme's decision (src/saddle_point_gpu.jl:419-622) on
    scripted weighted-KKT errors, one case per branch, expected values read
    off the reference: the artificial restart (restart length at least 0.36
    times the iterations completed, `iteration` - 1), KKT_GREEDY (restart
    to the average iff current >= average) against NO_RESTART_TO_CURRENT
    (always the average), the adaptive test (ratio below 0.2; below 0.8
    and above the last trial; the last trial always updated), the fixed
    frequency and NO_RESTARTS; also which errors the reference computes."""
    NO, RESET, AVG = (RestartChoice.RESTART_CHOICE_NO_RESTART,
                      RestartChoice.RESTART_CHOICE_WEIGHTED_AVERAGE_RESET,
                      RestartChoice.RESTART_CHOICE_RESTART_TO_AVERAGE)

    def rp(scheme=2, metric=1):
        return construct_restart_parameters(scheme, metric, 1000, 0.36, 0.2, 0.8, 0.5)

    # (params, restart length, iteration, cur, avg, last, last trial) ->
    # (choice, last trial after, errors consulted)
    cases = [
        (rp(), 0, 100, 1.0, 0.5, 1.0, 0.7, NO, 0.7, set()),
        (rp(), 50, 100, 1.0, 0.5, 1.0, 0.7, AVG, 0.7, {"cur", "avg"}),    # 50 >= 35.64
        (rp(), 50, 100, 0.4, 0.5, 1.0, 0.7, RESET, 0.7, {"cur", "avg"}),
        (rp(), 36, 101, 0.4, 0.5, 1.0, 0.7, RESET, 0.7, {"cur", "avg"}),  # 36 >= 36.0
        (rp(), 35, 99, 1.0, 0.5, 1.0, 0.7, NO, 0.5, {"cur", "avg", "last"}),  # 35 < 35.28; 0.5 < 0.7
        (rp(), 10, 100, 1.0, 0.1, 1.0, 0.7, AVG, 0.1, {"cur", "avg", "last"}),   # 0.1 < 0.2
        (rp(), 10, 100, 1.0, 0.5, 1.0, 0.4, AVG, 0.5, {"cur", "avg", "last"}),   # 0.5 > 0.4
        (rp(), 10, 100, 1.0, 0.5, 1.0, 0.9, NO, 0.5, {"cur", "avg", "last"}),    # 0.5 < 0.9
        (rp(), 10, 100, 1.0, 0.9, 1.0, 0.1, NO, 0.9, {"cur", "avg", "last"}),    # 0.9 >= 0.8
        (rp(), 10, 100, 0.3, 0.5, 2.0, 0.1, RESET, 0.15, {"cur", "avg", "last"}),  # current: 0.15
        (rp(metric=0), 10, 100, 0.1, 0.5, 1.0, 0.4, AVG, 0.5, {"cur", "avg", "last"}),
        (rp(scheme=1), 1000, 10000, 1.0, 2.0, 1.0, 0.7, RESET, 0.7, {"cur", "avg"}),
        (rp(scheme=1), 999, 10000, 1.0, 0.5, 1.0, 0.7, NO, 0.7, {"cur", "avg"}),
        (rp(scheme=0), 10, 100, 1.0, 0.5, 1.0, 0.7, NO, 0.7, set()),
        (rp(scheme=0), 50, 100, 1.0, 0.5, 1.0, 0.7, RESET, 0.7, set()),
    ]
    for k, (p, length, it, cur, avg, last, trial, want, trial_after, asked) in enumerate(cases):
        got, t, a = O.restart_choice(p, length, it, cur, avg, last, trial)
        assert (got, t, a) == (want, trial_after, asked), (k, got, t, a)


def test_primal_weight_update():
    """compute_new_primal_weight (src/saddle_point_gpu.jl:627-654): a
    log-space average of the old weight and dual / primal distance with
    weight `smoothing`; no change unless both distances exceed eps()."""
    eps = 2.220446049250313e-16
    assert math.isclose(O.primal_weight(1.0, 4.0, 1.0, 0.5), 2.0, rel_tol=1e-15)
    assert math.isclose(O.primal_weight(2.0, 6.0, 5.0, 1.0), 3.0, rel_tol=1e-15)
    assert math.isclose(O.primal_weight(2.0, 6.0, 5.0, 0.0), 5.0, rel_tol=1e-15)
    assert math.isclose(O.primal_weight(1.0, 100.0, 0.01, 0.5), 1.0, rel_tol=1e-15)
    assert math.isclose(O.primal_weight(2.3e-16, 2.3e-16, 4.0, 0.5), 2.0, rel_tol=1e-15)
    for p, d in ((1e-17, 1.0), (1.0, 0.0), (eps, 1.0), (1.0, eps)):
        assert O.primal_weight(p, d, 3.7, 0.5) == 3.7


def test_fixed_frequency_and_no_restarts_run():
    prob, _ = config_lp(0, scale=0.25)
    rp = construct_restart_parameters(1, 1, 64, 1.0, 0.2, 0.8, 0.5)
    out = O.solve(prob, params(1e-6, restart_params=rp))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    rp = construct_restart_parameters(0, 1, 1000, 1.0, 0.2, 0.8, 0.5)
    out = O.solve(prob, params(1e-4, restart_params=rp))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL


def test_invalid_inputs_raise():
    prob, _ = lp_cases.tiny_optimal()
    prob.objective_vector[0] = math.nan
    with pytest.raises(O.OracleError):
        O.solve(prob, params())
    prob, _ = lp_cases.tiny_optimal()
    prob.variable_lower_bound[0] = math.inf
    with pytest.raises(O.OracleError):
        O.solve(prob, params())
    prob, _ = lp_cases.tiny_optimal()
    with pytest.raises(O.OracleError):
        O.solve(prob, params(primal_importance=-1.0))


def test_malformed_colptr_is_refused_before_row_reads():
    """colptr = [1, 10, 3] with nnz = 2: the endpoints agree with nnz, the
    middle entry points past rowval; validation rejects it before reading
    any row index."""
    import ctypes as C
    from cupdlp import _abi
    from cupdlp.solver import to_c_params
    for colptr in ([1, 10, 3], [1, 3, 2, 3]):
        lp, keep = lp_cases.raw_lp(colptr, [1, 2], [1.0, 1.0], m=2)
        res = _abi.pdlp_result()
        x, y = np.zeros(len(colptr) - 1), np.zeros(2)
        res.x = x.ctypes.data_as(C.POINTER(C.c_double))
        res.y = y.ctypes.data_as(C.POINTER(C.c_double))
        cp = to_c_params(params())
        rc = O.lib().orc_solve(C.byref(lp), C.byref(cp), C.byref(res), _abi.STATS_CB(), None)
        assert rc > 0
        assert b"colptr" in O.lib().orc_last_error()


# ---------------------------------------------------------------- rescaling

def _dense_scaled(prob, vals):
    A = prob.constraint_matrix.copy()
    A.data = vals.copy()
    return A.toarray()


def test_ruiz_diagonal_closed_form():
    # diag(4, 9): one Ruiz pass divides by sqrt(4)*sqrt(4), sqrt(9)*sqrt(9) -> identity;
    # Pock-Chambolle then sees unit rows/columns.  E = D = (2, 3).
    prob = lp_cases.lp([[4.0, 0.0], [0.0, 9.0]], [1.0, 1.0], [1.0, 1.0], [0, 0], [10, 10], 2)
    vals, c, l, u, b, E, D = O.rescale(prob, PdhgParameters())
    np.testing.assert_array_equal(vals, [1.0, 1.0])
    np.testing.assert_array_equal(E, [2.0, 3.0])
    np.testing.assert_array_equal(D, [2.0, 3.0])
    np.testing.assert_array_equal(c, [0.5, 1.0 / 3.0])
    np.testing.assert_array_equal(u, [20.0, 30.0])
    np.testing.assert_array_equal(b, [0.5, 1.0 / 3.0])


def test_rescaling_is_a_diagonal_similarity():
    prob, _ = random_lp(n=80, m=40, seed=3, row_scale_decades=2.0, col_scale_decades=2.0)
    vals, c, l, u, b, E, D = O.rescale(prob, PdhgParameters())
    A = prob.constraint_matrix.toarray()
    At = _dense_scaled(prob, vals)
    np.testing.assert_allclose(At, A / E[:, None] / D[None, :], rtol=1e-13)
    np.testing.assert_allclose(c, prob.objective_vector / D, rtol=1e-14)
    np.testing.assert_allclose(b, prob.right_hand_side / E, rtol=1e-14)
    fin = np.isfinite(prob.variable_upper_bound)
    np.testing.assert_allclose(u[fin], prob.variable_upper_bound[fin] * D[fin], rtol=1e-14)


def test_pock_chambolle_operator_norm_bound():
    # Pock & Chambolle (2011), Lemma 2: with alpha = 1 the scaled matrix has ||A||_2 <= 1.
    prob, _ = random_lp(n=120, m=60, seed=4, row_scale_decades=3.0, col_scale_decades=2.0)
    for ruiz in (0, 10):
        vals, *_ = O.rescale(prob, PdhgParameters(l_inf_ruiz_iterations=ruiz))
        s = np.linalg.norm(_dense_scaled(prob, vals), 2)
        assert s <= 1.0 + 1e-12


def test_ruiz_equilibrates_infinity_norms():
    prob, _ = random_lp(n=60, m=30, seed=5, row_scale_decades=3.0, col_scale_decades=2.0)
    vals, *_ = O.rescale(prob, PdhgParameters(l_inf_ruiz_iterations=40, pock_chambolle_alpha=None))
    At = np.abs(_dense_scaled(prob, vals))
    np.testing.assert_allclose(At.max(axis=0), 1.0, atol=1e-6)
    np.testing.assert_allclose(At.max(axis=1), 1.0, atol=1e-6)


# ---------------------------------------------------------------- KKT statistics

def test_kkt_stats_vanish_at_certificate():
    prob, cert = config_lp(0)
    ci, ii = O.point_stats(prob, cert.x, cert.y)
    scale = 1 + abs(cert.objective)
    assert ci.l2_primal_residual < 1e-10 * (1 + np.linalg.norm(prob.right_hand_side))
    assert ci.l2_dual_residual < 1e-10 * (1 + np.linalg.norm(prob.objective_vector))
    assert abs(ci.primal_objective - ci.dual_objective) < 1e-10 * scale
    assert abs(ci.primal_objective - cert.objective) < 1e-12 * scale


def test_kkt_stats_closed_form_perturbations():
    # A = I (2 rows: one equality, one >=), bounds [0, 1]; b = (0.5, 0.25), c = (1, 1).
    prob = lp_cases.lp(np.eye(2), [0.5, 0.25], [1.0, 1.0], [0, 0], [1, 1], 1)
    x = np.array([0.7, 1.5])          # eq row residual 0.5-0.7=-0.2; >= row satisfied; x2 above u by 0.5
    y = np.array([2.0, -0.3])         # >= row dual negative: dual residual 0.3
    ci, ii = O.point_stats(prob, x, y, 1e-6, 1e-6)
    assert math.isclose(ci.l_inf_primal_residual, 0.5, rel_tol=0, abs_tol=1e-15)
    assert math.isclose(ci.l2_primal_residual, math.hypot(0.2, 0.5), rel_tol=1e-15)
    assert math.isclose(ci.primal_objective, 2.2, rel_tol=1e-15)
    # gradient g = c - A^T y = (-1, 1.3): reduced costs r = (min(g1,0), max(g2,0)) = (-1, 1.3),
    # no violation; dual residual only from y2 < 0.
    assert math.isclose(ci.l2_dual_residual, 0.3, rel_tol=1e-15)
    # dual objective = b'y + sum(l r+ + u r-) = (1 - 0.075) + (0*1.3 + 1*(-1)) = -0.075
    assert math.isclose(ci.dual_objective, -0.075, rel_tol=1e-12)
    assert math.isclose(ci.l_inf_primal_variable, 1.5, rel_tol=0)


# ---------------------------------------------------------------- step-size rule

def test_step_rule_reject_and_shrink():
    # eta = 1 > limit = movement/interaction = 0.5 -> reject; attempt k = 1:
    # new eta = min((1 - 1/2^0.3) * 0.5, (1 + 1/2^0.6) * 1), written as the reference
    # does (pdhg:405-409: 1/(k+1)^r, not (k+1)^-r, which rounds differently)
    eta, acc, ne = O.step_rule(1.0, -1.0, 0.5, 1)
    assert not acc and not ne
    assert eta == min((1 - 1 / 2.0 ** 0.3) * 0.5, (1 + 1 / 2.0 ** 0.6) * 1.0)


def test_step_rule_accept_and_grow():
    eta, acc, ne = O.step_rule(0.1, 1.0, 0.5, 9)
    assert acc and not ne
    assert eta == min((1 - 1 / 10.0 ** 0.3) * 0.5, (1 + 1 / 10.0 ** 0.6) * 0.1)


def test_step_rule_zero_interaction_and_numerical_error():
    eta, acc, ne = O.step_rule(0.25, 0.0, 3.0, 4)
    assert acc and eta == (1 + 1 / 5.0 ** 0.6) * 0.25      # limit = Inf
    eta, acc, ne = O.step_rule(0.25, 1e-3, 0.0, 4)
    assert ne and not acc and eta == 0.25


# ---------------------------------------------------------------- SpMV

def test_spmv_matches_dense():
    rng = np.random.default_rng(9)
    A = sp.random(37, 53, density=0.2, random_state=3, format="csc")
    prob = lp_cases.lp(A.toarray(), np.zeros(37), np.zeros(53), np.zeros(53), np.ones(53), 10)
    x, y = rng.standard_normal(53), rng.standard_normal(37)
    Ax, ATy = O.spmv(prob, x, y)
    np.testing.assert_allclose(Ax, A @ x, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(ATy, A.T @ y, rtol=1e-13, atol=1e-14)
