"""GPU parity of the branches the default solve does not take, bitwise
against the CPU oracle (the bar of tests/test_gpu_parity.py):

* l2_norm_rescaling (src/preprocess.jl:89-103, 344-358; host rescaling path);
* the L_INF optimality norm (src/termination.jl:142-170);
* the KKT-pass and time limits (src/termination.jl:236-245);
* NUMERICAL_ERROR, reached through the movement = 0 fault injection
  (src/primal_dual_hybrid_gradient_gpu.jl:384-390, 684-686);
* a malformed colptr refused before any row index is read;
* FP32 + refinement within the caller's iteration limit, through
  pdlp_solve, pdlp_batch_solve and the request loop.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import lp_cases
import oracle as O
from cupdlp import (OptimalityNorm, PdhgParameters, Precision, TerminationReason as TR,
                    construct_termination_criteria)
from cupdlp.generators import config_lp, random_lp, transportation_lp
from test_gpu_parity import compare_runs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params(tol=1e-12, iteration_limit=2**31 - 1, norm=OptimalityNorm.L2, kkt_limit=math.inf,
           time_limit=math.inf, **kw):
    tc = construct_termination_criteria(optimality_norm=norm, eps_optimal_absolute=tol,
                                        eps_optimal_relative=tol, iteration_limit=iteration_limit,
                                        kkt_matrix_pass_limit=kkt_limit, time_sec_limit=time_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def test_l2_norm_rescaling_bitwise(gpu_ctx):
    say("case: l2_norm_rescaling = true (n=2000 LP to 1e-8; scaled Zipf LP n=20000, 300 iterations)")
    prob, cert = config_lp(0)
    p = params(1e-8, l2_norm_rescaling=True)
    g, o = gpu_ctx.solve(prob, p), O.solve(prob, p)
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    prob, _ = config_lp(1, scale=0.02)
    p = params(iteration_limit=300, l2_norm_rescaling=True)
    compare_runs(gpu_ctx.solve(prob, p), O.solve(prob, p))
    # the L2 pass alone (no Ruiz, no Pock-Chambolle)
    p = params(iteration_limit=300, l2_norm_rescaling=True, l_inf_ruiz_iterations=0,
               pock_chambolle_alpha=None)
    compare_runs(gpu_ctx.solve(prob, p), O.solve(prob, p))


def test_l_inf_optimality_norm_bitwise(gpu_ctx):
    say("case: optimality_norm = L_INF (n=2000 LP and a 40x50 transportation LP to 1e-6)")
    for prob in (config_lp(0)[0], transportation_lp(40, 50, seed=6)):
        p = params(1e-6, norm=OptimalityNorm.L_INF, termination_evaluation_frequency=8)
        g, o = gpu_ctx.solve(prob, p), O.solve(prob, p)
        compare_runs(g, o)
        assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL


def test_kkt_pass_limit_bitwise(gpu_ctx):
    say("case: kkt_matrix_pass_limit at 2.5, 40.0 and 300.0")
    prob, _ = config_lp(0, scale=0.5)
    for lim in (2.5, 40.0, 300.0):
        p = params(kkt_limit=lim)
        g, o = gpu_ctx.solve(prob, p), O.solve(prob, p)
        compare_runs(g, o)
        assert g.termination_reason == TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT
        assert g.iteration_stats[-1].cumulative_kkt_matrix_passes >= lim
    g = gpu_ctx.solve(prob, params(kkt_limit=2.5))
    assert g.iteration_count == 0  # 0.5 for the initial norm + 2 for the first evaluation


def test_time_limit_zero(gpu_ctx):
    say("case: time_sec_limit = 0 (stops at the first evaluation), precedence of the limits")
    prob, _ = config_lp(0, scale=0.5)
    g = gpu_ctx.solve(prob, params(time_limit=0.0))
    o = O.solve(prob, params(time_limit=0.0))
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_TIME_LIMIT and g.iteration_count == 0
    g = gpu_ctx.solve(prob, params(kkt_limit=1.0, time_limit=0.0))
    assert g.termination_reason == TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT
    g = gpu_ctx.solve(prob, params(iteration_limit=0, kkt_limit=1.0, time_limit=0.0))
    assert g.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT


@pytest.mark.parametrize("attempt", [1, 5, 23, 200])
def test_numerical_error_bitwise(gpu_ctx, attempt):
    say(f"case: NUMERICAL_ERROR injected at attempt {attempt}")
    prob, _ = config_lp(0, scale=0.5)
    p = params(debug_numerical_error_attempt=attempt)
    g, o = gpu_ctx.solve(prob, p, want_current=True), O.solve(prob, p, want_current=True)
    compare_runs(g, o)
    assert g.termination_reason == TR.TERMINATION_REASON_NUMERICAL_ERROR
    assert g.total_attempts == attempt
    np.testing.assert_array_equal(g.primal_current, o.primal_current)
    # in a batch every instance shares the parameters, so each stops at its
    # own attempt `attempt`, exactly as when solved alone
    others = [random_lp(n=n, m=n // 2, seed=s)[0] for n, s in ((400, 41), (3000, 42))]
    outs = gpu_ctx.batch_solve([prob] + others, p)
    for q, b in zip([prob] + others, outs):
        a = gpu_ctx.solve(q, p)
        assert b.termination_reason == a.termination_reason
        assert b.iteration_count == a.iteration_count
        assert np.array_equal(a.primal_solution, b.primal_solution)


def test_malformed_colptr_refused(gpu_ctx):
    say("case: colptr = [1, 10, 3] with nnz = 2 (refused by validation)")
    from cupdlp import _abi
    from cupdlp.solver import to_c_params
    lib = gpu_ctx.lib
    for colptr in ([1, 10, 3], [1, 3, 2, 3]):
        lp, keep = lp_cases.raw_lp(colptr, [1, 2], [1.0, 1.0], m=2)
        x, y = np.zeros(len(colptr) - 1), np.zeros(2)
        res = _abi.pdlp_result()
        res.x = x.ctypes.data_as(C.POINTER(C.c_double))
        res.y = y.ctypes.data_as(C.POINTER(C.c_double))
        cp = to_c_params(params())
        rc = lib.pdlp_solve(gpu_ctx.handle, C.byref(lp), C.byref(cp), C.byref(res),
                            _abi.STATS_CB(), None)
        assert rc > 0
        assert "colptr" in gpu_ctx.error()


def test_fp32_first_solve_cap_falls_back(gpu_ctx):
    """An inner tolerance FP32 data cannot reach (1e-13) stops the first FP32
    solve at its 100,000-iteration cap; the refinement counts as stalled and
    the FP64 solve finishes the LP, bitwise as a plain FP64 solve."""
    say("case: FP32 + refinement with refine_inner_tolerance 1e-13 (first-solve cap)")
    prob, _ = config_lp(0, scale=0.5)
    g = gpu_ctx.solve(prob, params(1e-8, precision=Precision.FP32_REFINE,
                                   refine_inner_tolerance=1e-13))
    f = gpu_ctx.solve(prob, params(1e-8))
    assert g.fp64_fallback and g.refinement_rounds == 0
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    # the cap acts at the first evaluation (every 64 iterations) at or past it
    assert 100000 <= g.iteration_count - f.iteration_count < 100064
    assert np.array_equal(g.primal_solution, f.primal_solution)
    assert np.array_equal(g.dual_solution, f.dual_solution)


def test_fp32_refine_stays_within_iteration_limit(gpu_ctx):
    """The default correction-solve cap (max(10000, 20 x round 0)) is far above
    the limit: the refinement rounds together must stop at it."""
    say("case: FP32 + refinement, iteration_limit 300 and 1000, default inner cap")
    from cupdlp.solver import Server
    probs = [config_lp(0)[0], random_lp(n=5000, m=2500, seed=44)[0],
             transportation_lp(40, 50, seed=6)]
    for lim in (300, 1000):
        p = params(1e-10, iteration_limit=lim, precision=Precision.FP32_REFINE)
        for prob in probs:
            g = gpu_ctx.solve(prob, p)
            assert g.iteration_count <= lim, (lim, g.iteration_count)
            if g.termination_reason != TR.TERMINATION_REASON_OPTIMAL:
                assert g.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
        for g in gpu_ctx.batch_solve(probs, p):
            assert g.iteration_count <= lim
        srv = Server(gpu_ctx, p)
        ids = [srv.submit(q, 1e-10) for q in probs]
        got = {}
        for _ in range(100000):
            for f in srv.poll():
                got[f] = srv.result(f)
            if len(got) == len(ids):
                break
        srv.close()
        assert len(got) == len(ids)
        for r in got.values():
            assert r.iteration_count <= lim
