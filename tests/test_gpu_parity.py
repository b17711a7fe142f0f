"""GPU parity: libpdlp_hip.so (through the C ABI) against the CPU oracle.

Bar (FP64 path): bit-for-bit.  Both sides compute every sum (SpMV rows,
dot products, norms) as a compensated sum of error-free products rounded
once (Dot2: error <= eps |sum| + gamma_n^2 sum |terms|), so the two
summation orders give the same FP64 value unless a sum cancels to ~n^2 eps
of its terms; every elementwise formula is evaluated in the same written
order without contraction.  Step sizes, primal weights, restart decisions, every
recorded KKT statistic and the returned iterates are then equal, and the
tests ask for that.  (Bitwise equality is stronger than the required
1e-9 relative bar over the first 1,000 iterations; a heavily cancelling sum
rounding differently in the two orders would show up here as a last-bit
difference, to be judged against that 1e-9 bar.)
"""
import math
import sys

import numpy as np
import pytest

import lp_cases
import oracle as O
from cupdlp import PdhgParameters, TerminationReason as TR, construct_termination_criteria
from cupdlp.generators import config_lp, random_lp, transportation_lp

pytestmark = pytest.mark.gpu

REL = 0.0  # bitwise


def params(tol=1e-4, iteration_limit=2**31 - 1, **kw):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def vec_rel(a, b):
    na = max(np.linalg.norm(a), np.linalg.norm(b), 1e-300)
    return np.linalg.norm(a - b) / na


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CI_FIELDS = ("primal_objective", "dual_objective", "corrected_dual_objective",
             "l_inf_primal_residual", "l2_primal_residual", "l_inf_dual_residual",
             "l2_dual_residual", "relative_l_inf_primal_residual", "relative_l2_primal_residual",
             "relative_l_inf_dual_residual", "relative_l2_dual_residual",
             "relative_optimality_gap", "l_inf_primal_variable", "l2_primal_variable",
             "l_inf_dual_variable", "l2_dual_variable")
II_FIELDS = ("max_primal_ray_infeasibility", "primal_ray_linear_objective",
             "max_dual_ray_infeasibility", "dual_ray_objective")


def same(a, b, tol):
    if tol == 0.0:
        return a == b or (math.isnan(a) and math.isnan(b))
    return rel(a, b) <= tol


def compare_runs(g, o, rel_tol=REL, vec_tol=REL, check_all_stats=True):
    if check_all_stats:
        assert len(g.iteration_stats) == len(o.iteration_stats)
        for sg, so in zip(g.iteration_stats, o.iteration_stats):
            it = sg.iteration_number
            assert it == so.iteration_number
            assert sg.restart_used == so.restart_used, (it, sg.restart_used, so.restart_used)
            assert same(sg.step_size, so.step_size, rel_tol), (it, sg.step_size, so.step_size)
            assert same(sg.primal_weight, so.primal_weight, rel_tol), (it, sg.primal_weight, so.primal_weight)
            assert sg.cumulative_kkt_matrix_passes == so.cumulative_kkt_matrix_passes
            cg, co = sg.convergence_information[0], so.convergence_information[0]
            for f in CI_FIELDS:
                assert same(getattr(cg, f), getattr(co, f), rel_tol), (it, f, getattr(cg, f), getattr(co, f))
            ig, io = sg.infeasibility_information[0], so.infeasibility_information[0]
            for f in II_FIELDS:
                assert same(getattr(ig, f), getattr(io, f), rel_tol), (it, f, getattr(ig, f), getattr(io, f))
    assert g.termination_reason == o.termination_reason, (g.termination_reason, o.termination_reason)
    assert g.iteration_count == o.iteration_count
    assert g.total_attempts == o.total_attempts
    if vec_tol == 0.0:
        assert np.array_equal(g.primal_solution, o.primal_solution)
        assert np.array_equal(g.dual_solution, o.dual_solution)
    else:
        assert vec_rel(g.primal_solution, o.primal_solution) <= vec_tol
        assert vec_rel(g.dual_solution, o.dual_solution) <= vec_tol


def test_c0_first_1000_iterations(gpu_ctx):
    say("case: n=2000 LP seed 0, iteration_limit=1000, FP64")
    prob, cert = config_lp(0)
    p = params(tol=1e-12, iteration_limit=1000)
    g = gpu_ctx.solve(prob, p, want_current=True)
    o = O.solve(prob, p, want_current=True)
    assert g.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    compare_runs(g, o)
    assert np.array_equal(g.primal_current, o.primal_current)
    assert np.array_equal(g.dual_current, o.dual_current)


def test_host_and_device_preconditioning_agree(gpu_ctx):
    say("case: n=2000 LP, rescaling on the host vs on the device")
    prob, _ = config_lp(0)
    a = gpu_ctx.solve(prob, params(1e-8, device_preconditioning=True))
    b = gpu_ctx.solve(prob, params(1e-8, device_preconditioning=False))
    compare_runs(a, b)
    prob, _ = config_lp(1, scale=0.01)
    a = gpu_ctx.solve(prob, params(1e-12, iteration_limit=300, device_preconditioning=True))
    b = gpu_ctx.solve(prob, params(1e-12, iteration_limit=300, device_preconditioning=False))
    compare_runs(a, b)


@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_c0_to_tolerance(gpu_ctx, tol):
    say(f"case: n=2000 LP seed 0 to {tol}")
    prob, cert = config_lp(0)
    p = params(tol=tol)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    ci = g.iteration_stats[-1].convergence_information[0]
    # reference optimality test (src/termination.jl:142-170) holds on the output
    assert abs(ci.primal_objective - cert.objective) <= 50 * tol * (1 + abs(cert.objective))


def test_closed_forms(gpu_ctx):
    say("case: closed-form tiny LPs")
    prob, opt = lp_cases.tiny_optimal()
    g = gpu_ctx.solve(prob, params(1e-8))
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert abs(g.primal_solution[0] - 1.0) < 1e-6
    assert abs(g.iteration_stats[-1].convergence_information[0].primal_objective - opt) < 1e-6
    prob, opt = lp_cases.tiny_two_var()
    g = gpu_ctx.solve(prob, params(1e-8))
    o = O.solve(prob, params(1e-8))
    compare_runs(g, o)
    assert abs(g.iteration_stats[-1].convergence_information[0].primal_objective - opt) < 1e-6


def test_infeasible_and_unbounded(gpu_ctx):
    say("case: primal infeasible / dual infeasible")
    g = gpu_ctx.solve(lp_cases.primal_infeasible(), params(1e-6))
    o = O.solve(lp_cases.primal_infeasible(), params(1e-6))
    assert g.termination_reason == o.termination_reason == TR.TERMINATION_REASON_PRIMAL_INFEASIBLE
    assert g.iteration_count == o.iteration_count >= 9  # detection allowed from iteration 10
    g = gpu_ctx.solve(lp_cases.dual_infeasible(), params(1e-6))
    o = O.solve(lp_cases.dual_infeasible(), params(1e-6))
    assert g.termination_reason == o.termination_reason == TR.TERMINATION_REASON_DUAL_INFEASIBLE
    assert g.iteration_count == o.iteration_count


def test_no_rows(gpu_ctx):
    say("case: m = 0")
    prob, opt = lp_cases.no_rows()
    g = gpu_ctx.solve(prob, params(1e-8))
    o = O.solve(prob, params(1e-8))
    assert g.termination_reason == o.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    np.testing.assert_allclose(g.primal_solution, [0.0, 2.0], atol=1e-6)
    assert abs(g.iteration_stats[-1].convergence_information[0].primal_objective - opt) < 1e-6


def test_empty_row_and_column(gpu_ctx):
    say("case: an empty equality row, an empty >= row and an empty column, no presolve; "
        "FP64 bitwise vs the oracle, FP32 + refinement to the certificate")
    from cupdlp import Precision
    prob, obj = lp_cases.empty_row_and_column_lp()
    g = gpu_ctx.solve(prob, params(1e-8))
    o = O.solve(prob, params(1e-8))
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert g.primal_solution[-1] == 0.0
    assert abs(g.iteration_stats[-1].convergence_information[0].primal_objective - obj) \
        <= 1e-6 * (1 + abs(obj))
    f = gpu_ctx.solve(prob, params(1e-8, precision=Precision.FP32_REFINE))
    assert f.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert abs(f.iteration_stats[-1].convergence_information[0].primal_objective - obj) \
        <= 1e-6 * (1 + abs(obj))


def test_inequality_only_and_fixed_variables(gpu_ctx):
    say("case: an LP with no equality rows, one with a tenth of its variables fixed "
        "(lb = ub), one with one-sided bounds only, FP64 bitwise vs the oracle")
    for prob in (random_lp(n=300, m=150, num_eq=0, seed=41)[0],
                 lp_cases.fixed_variables_lp()[0], lp_cases.one_sided_bounds_lp()[0]):
        g = gpu_ctx.solve(prob, params(1e-8))
        o = O.solve(prob, params(1e-8))
        compare_runs(g, o)
        assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL


def test_iteration_limit_semantics(gpu_ctx):
    say("case: iteration_limit = 37")
    prob, _ = config_lp(0, scale=0.25)
    g = gpu_ctx.solve(prob, params(1e-12, iteration_limit=37))
    o = O.solve(prob, params(1e-12, iteration_limit=37))
    assert g.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    assert g.iteration_count == o.iteration_count == 37
    compare_runs(g, o)


def test_long_rows(gpu_ctx):
    say("case: LP with a 6000-entry row (chunked long-row path)")
    prob, opt = lp_cases.long_row_lp(6000)
    p = params(1e-12, iteration_limit=500)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)


def test_very_long_row_chunked(gpu_ctx):
    say("case: LP with a 140000-entry row (split into chunks combined across work-groups)")
    prob, opt = lp_cases.long_row_lp(140000, seed=8)
    p = params(1e-12, iteration_limit=200)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)


def test_long_column(gpu_ctx):
    say("case: LP with a 20000-entry column (chunked long rows of A^T in the primal step)")
    prob, opt = lp_cases.long_col_lp(20000)
    p = params(1e-12, iteration_limit=500)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)


def test_transportation_small(gpu_ctx):
    say("case: 60 x 60 transportation LP to 1e-6 (has rejected step attempts)")
    prob = transportation_lp(60, 60, seed=5)
    p = params(1e-6)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    assert o.total_attempts > o.iteration_count  # exercises rejection after acceptance
    compare_runs(g, o)


def test_scaled_zipf_lp(gpu_ctx):
    say("case: n=20000 Zipf-degree scaled LP, 640 iterations")
    prob, cert = config_lp(1, scale=0.02)
    p = params(1e-12, iteration_limit=640)
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)


def test_batch_alone_vs_mixed_bitwise(gpu_ctx):
    say("case: batch of 4 vs alone")
    probs = [random_lp(n=n, m=n // 2, seed=s)[0] for n, s in ((300, 11), (1200, 12), (5000, 13), (700, 14))]
    p = params(1e-6)
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    for a, b in zip(alone, mixed):
        assert a.termination_reason == b.termination_reason
        assert a.iteration_count == b.iteration_count
        assert np.array_equal(a.primal_solution, b.primal_solution)
        assert np.array_equal(a.dual_solution, b.dual_solution)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_of_degenerate_shapes_alone_vs_mixed(gpu_ctx, precision):
    say(f"case: batch of 9 degenerate / structured LPs vs alone, {precision} (empty rows "
        "and column, fixed and one-sided bounds, no equalities, long column, long row, "
        "transportation, m = 0)")
    from cupdlp import Precision
    probs = [lp_cases.empty_row_and_column_lp()[0], lp_cases.fixed_variables_lp()[0],
             lp_cases.one_sided_bounds_lp()[0], random_lp(n=300, m=150, num_eq=0, seed=41)[0],
             lp_cases.long_col_lp(3000)[0], lp_cases.long_row_lp(6000)[0],
             transportation_lp(30, 40, seed=4), lp_cases.no_rows()[0],
             random_lp(n=900, m=450, seed=15)[0]]
    prec = Precision.FP64 if precision == "fp64" else Precision.FP32_REFINE
    p = params(1e-6, precision=prec)
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert a.termination_reason == b.termination_reason == TR.TERMINATION_REASON_OPTIMAL, k
        assert a.iteration_count == b.iteration_count, k
        assert np.array_equal(a.primal_solution, b.primal_solution), k
        assert np.array_equal(a.dual_solution, b.dual_solution), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_infeasible_and_unbounded_in_a_batch(gpu_ctx, precision):
    say(f"case: primal infeasible and unbounded random LPs ({precision}; FP64 bitwise vs the "
        "oracle), then alone vs a batch with tiny infeasible LPs and optimal ones")
    from cupdlp import Precision
    infeas, unbdd = lp_cases.infeasible_random_lp(), lp_cases.unbounded_random_lp()
    p = params(1e-6, precision=Precision.FP64 if precision == "fp64" else Precision.FP32_REFINE)
    if precision == "fp64":
        for prob in (infeas, unbdd):
            compare_runs(gpu_ctx.solve(prob, p), O.solve(prob, p))
    probs = [infeas, random_lp(n=600, m=300, seed=16)[0], unbdd, lp_cases.primal_infeasible(),
             lp_cases.dual_infeasible(), random_lp(n=2000, m=1000, seed=17)[0]]
    want = [TR.TERMINATION_REASON_PRIMAL_INFEASIBLE, TR.TERMINATION_REASON_OPTIMAL,
            TR.TERMINATION_REASON_DUAL_INFEASIBLE, TR.TERMINATION_REASON_PRIMAL_INFEASIBLE,
            TR.TERMINATION_REASON_DUAL_INFEASIBLE, TR.TERMINATION_REASON_OPTIMAL]
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert a.termination_reason == b.termination_reason == want[k], k
        assert a.iteration_count == b.iteration_count, k
        assert np.array_equal(a.primal_solution, b.primal_solution), k
        assert np.array_equal(a.dual_solution, b.dual_solution), k


def test_constant_step_policy(gpu_ctx):
    """ConstantStepsizeParams: step (1 - 0.2) / sigma_max from the power method
    (start vector SplitMix64 + Box-Muller on both sides), every attempt taken;
    bitwise equal to the oracle, with device and with host rescaling."""
    from cupdlp import ConstantStepsizeParams
    say("case: n=2000 LP and a transportation LP, constant step policy, 600 iterations")
    for prob in (config_lp(0)[0], transportation_lp(30, 40, seed=4)):
        for dev in (True, False):
            p = params(1e-12, iteration_limit=600, step_size_policy_params=ConstantStepsizeParams(),
                       device_preconditioning=dev)
            g = gpu_ctx.solve(prob, p)
            o = O.solve(prob, p)
            compare_runs(g, o)
            assert g.total_attempts == 0
    prob, opt = lp_cases.tiny_two_var()
    p = params(1e-8, step_size_policy_params=ConstantStepsizeParams())
    g = gpu_ctx.solve(prob, p)
    o = O.solve(prob, p)
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL


def test_invalid_input_raises(gpu_ctx):
    say("case: NaN objective")
    from cupdlp.solver import PdlpError
    prob, _ = lp_cases.tiny_optimal()
    prob.objective_vector[0] = math.nan
    with pytest.raises(PdlpError) as ei:
        gpu_ctx.solve(prob, params())
    assert ei.value.code > 0
