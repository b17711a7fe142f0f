"""pdlp_solve_noalloc (no device allocation, a fixed per-device region)
against pdlp_solve, through the C ABI: bitwise the same results on two
calls in a row, the second with a smaller LP; device free memory unchanged
by a call; oversized LPs and FP32 refused with their codes."""
import ctypes as C

import numpy as np
import pytest

from cupdlp import (PdhgParameters, PdlpError, Precision, TerminationReason as TR,
                    construct_termination_criteria)
from cupdlp.generators import config_lp, random_lp, transportation_lp

pytestmark = pytest.mark.gpu


def params(tol, iteration_limit=2**31 - 1, **kw):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def same(a, b):
    assert a.termination_reason == b.termination_reason
    assert a.iteration_count == b.iteration_count
    assert a.total_attempts == b.total_attempts
    assert np.array_equal(a.primal_solution, b.primal_solution)
    assert np.array_equal(a.dual_solution, b.dual_solution)
    assert len(a.iteration_stats) == len(b.iteration_stats)
    for sa, sb in zip(a.iteration_stats, b.iteration_stats):
        assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight
        ca, cb = sa.convergence_information[0], sb.convergence_information[0]
        assert ca.primal_objective == cb.primal_objective
        assert ca.l2_primal_residual == cb.l2_primal_residual
        assert ca.l2_dual_residual == cb.l2_dual_residual


def free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_matches_solve_on_two_calls(gpu_ctx):
    print("case: pdlp_solve_noalloc vs pdlp_solve: n=20000 Zipf LP, then a 40x50 "
          "transportation LP, then the n=2000 LP to 1e-8", flush=True)
    big, _ = config_lp(1, scale=0.02)         # n = 20,000, ~240k nonzeros
    small = transportation_lp(40, 50, seed=6)  # n = 2,000
    p1 = params(1e-12, iteration_limit=700)
    p2 = params(1e-7)
    ref1, ref2 = gpu_ctx.solve(big, p1, want_current=True), gpu_ctx.solve(small, p2)
    f0 = free_device_bytes()
    a1 = gpu_ctx.solve_noalloc(big, p1, want_current=True)
    f1 = free_device_bytes()
    a2 = gpu_ctx.solve_noalloc(small, p2)  # second call, smaller input
    f2 = free_device_bytes()
    same(a1, ref1)
    assert np.array_equal(a1.primal_current, ref1.primal_current)
    same(a2, ref2)
    assert a2.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    # no device memory taken by the calls
    assert f0 == f1 == f2, (f0, f1, f2)
    # the region keeps the last call's contents: a third call still agrees
    prob, _ = config_lp(0)
    same(gpu_ctx.solve_noalloc(prob, params(1e-8)), gpu_ctx.solve(prob, params(1e-8)))


def test_noalloc_degenerate_shapes(gpu_ctx):
    print("case: pdlp_solve_noalloc vs pdlp_solve on degenerate / structured LPs (empty "
          "rows and column, fixed and one-sided bounds, long column and row, m = 0)",
          flush=True)
    import lp_cases
    probs = [lp_cases.empty_row_and_column_lp()[0], lp_cases.fixed_variables_lp()[0],
             lp_cases.one_sided_bounds_lp()[0], lp_cases.long_col_lp(3000)[0],
             lp_cases.long_row_lp(6000)[0], lp_cases.no_rows()[0]]
    for k, prob in enumerate(probs):
        a = gpu_ctx.solve_noalloc(prob, params(1e-8))
        same(a, gpu_ctx.solve(prob, params(1e-8)))
        assert a.termination_reason == TR.TERMINATION_REASON_OPTIMAL, k


def test_noalloc_refusals(gpu_ctx):
    print("case: pdlp_solve_noalloc refuses FP32 + refinement and an LP beyond its region",
          flush=True)
    prob, _ = config_lp(0)
    with pytest.raises(PdlpError) as ei:
        gpu_ctx.solve_noalloc(prob, params(1e-6, precision=Precision.FP32_REFINE))
    assert ei.value.code == 2
    huge, _ = random_lp(n=400_000, m=200_000, seed=1)  # 3.2M nonzeros: beyond 128 MiB
    with pytest.raises(PdlpError) as ei:
        gpu_ctx.solve_noalloc(huge, params(1e-4, iteration_limit=10))
    assert ei.value.code == 7
    # the next call is unaffected
    same(gpu_ctx.solve_noalloc(prob, params(1e-6)), gpu_ctx.solve(prob, params(1e-6)))
