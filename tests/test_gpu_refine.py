"""FP32 + FP64 iterative refinement on the GPU.

The refined point is checked in FP64 by the CPU oracle's statistics of the
original problem (the reference's optimality test, src/termination.jl:142-170,
with eps_abs = eps_rel = tol) and against HiGHS objectives; batching must not
change any instance's result (bitwise).
"""
import json
import os
import sys

import numpy as np
import pytest

import lp_cases
import oracle as O
from cupdlp import (PdhgParameters, PdlpError, Precision, TerminationReason as TR,
                   construct_termination_criteria)
from cupdlp.generators import batch_sizes, config_lp, random_lp, transportation_lp

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "highs_objectives.json")) as f:
    HIGHS = json.load(f)["objectives"]


def params(tol, time_limit=120.0, **kw):
    # the time limit only guards the test run; every case here finishes in seconds
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        time_sec_limit=time_limit)
    return PdhgParameters(termination_criteria=tc, precision=Precision.FP32_REFINE, **kw)


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def check_kkt(prob, out, tol):
    ci, _ = O.point_stats(prob, out.primal_solution, out.dual_solution, tol, tol)
    b = np.linalg.norm(prob.right_hand_side)
    c = np.linalg.norm(prob.objective_vector)
    gap = abs(ci.primal_objective - ci.dual_objective)
    assert ci.l2_primal_residual < tol * (1 + b), ci.l2_primal_residual
    assert ci.l2_dual_residual < tol * (1 + c), ci.l2_dual_residual
    assert gap < tol * (1 + abs(ci.primal_objective) + abs(ci.dual_objective)), gap
    return ci


@pytest.mark.parametrize("tol", [1e-4, 1e-6, 1e-8])
def test_refined_c0(gpu_ctx, tol):
    say(f"case: n=2000 LP, FP32 + refinement to {tol}")
    prob, cert = config_lp(0)
    out = gpu_ctx.solve(prob, params(tol))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    ci = check_kkt(prob, out, tol)
    h = HIGHS["c0_seed0"]["objective"]
    assert abs(ci.primal_objective - h) <= 10 * tol * (1 + abs(h))
    say(f"  iterations {out.iteration_count}, refinement rounds {out.refinement_rounds}")


def test_refined_transportation(gpu_ctx):
    say("case: 60 x 60 transportation, FP32 + refinement to 1e-8")
    prob = transportation_lp(60, 60, seed=5)
    out = gpu_ctx.solve(prob, params(1e-8))
    say(f"  {out.termination_string}, iterations {out.iteration_count}, "
        f"refinement rounds {out.refinement_rounds}, {out.solve_time_sec:.2f} s")
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    ci = check_kkt(prob, out, 1e-8)
    h = HIGHS["transport_60x60_s5"]["objective"]
    assert abs(ci.primal_objective - h) <= 1e-7 * abs(h)
    assert out.refinement_rounds <= 3


def test_transportation_2000x2000_against_highs(gpu_ctx):
    say("case: 2,000 x 2,000 transportation LP (n = 4,000,000), FP32 + refinement to 1e-8")
    prob = config_lp(2)[0]
    out = gpu_ctx.solve(prob, params(1e-8))
    say(f"  {out.termination_string}, iterations {out.iteration_count}, "
        f"rounds {out.refinement_rounds}, {out.solve_time_sec:.2f} s")
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    h = HIGHS["transport_2000x2000_s2"]["objective"]
    fs = out.iteration_stats[-1].convergence_information[0]
    assert abs(fs.primal_objective - h) <= 1e-8 * abs(h)


def test_inner_iteration_cap_and_best_point(gpu_ctx):
    """Correction solves capped at 64 iterations under a total iteration
    limit of 400: the refinement stalls and the FP64 solve gets what is left
    of the limit, too little to converge.  The solve ends at the limit and
    returns a point no worse than the best refined one (the FP64 point only
    if it is better), and its final statistics describe that point."""
    say("case: n=2000 LP, FP32 + refinement, inner cap 64, iteration limit 400")
    prob, cert = config_lp(0)
    p = params(1e-8, refine_inner_iteration_limit=64, refine_max_outer=6)
    p.termination_criteria.iteration_limit = 400
    out = gpu_ctx.solve(prob, p)
    say(f"  {out.termination_string}, iterations {out.iteration_count}, "
        f"rounds {out.refinement_rounds}")
    assert out.termination_reason in (TR.TERMINATION_REASON_OPTIMAL,
                                      TR.TERMINATION_REASON_ITERATION_LIMIT)
    assert out.iteration_count <= 400
    assert out.refinement_rounds <= 6

    def kkt(s):
        c = s.convergence_information[0]
        return max(c.relative_l2_primal_residual, c.relative_l2_dual_residual,
                   c.relative_optimality_gap)

    # one recorded statistics entry per refinement round (round 0 included)
    refined = out.iteration_stats[:out.refinement_rounds + 1]
    assert kkt(out.iteration_stats[-1]) <= min(kkt(s) for s in refined)
    ci, _ = O.point_stats(prob, out.primal_solution, out.dual_solution, 1e-8, 1e-8)
    assert out.iteration_stats
    fs = out.iteration_stats[-1].convergence_information[0]
    for k in ("l2_primal_residual", "l2_dual_residual", "primal_objective", "dual_objective"):
        a, b = getattr(fs, k), getattr(ci, k)
        assert abs(a - b) <= 1e-9 * (1 + abs(b)), (k, a, b)


def test_stalled_refinement_falls_back_to_fp64(gpu_ctx):
    """Correction solves capped at 64 iterations stall the refinement; the LP
    is then solved again in FP64, whose result is returned (bitwise the plain
    FP64 solve's), alone and in a batch."""
    say("case: n=2000 LP and a transportation LP, FP32 + refinement with inner cap 64 -> FP64 fallback")
    probs = [config_lp(0)[0], transportation_lp(40, 50, seed=6)]
    p = params(1e-8, refine_inner_iteration_limit=64)
    p64 = params(1e-8)
    p64.precision = Precision.FP64
    plain = [gpu_ctx.solve(q, p64) for q in probs]
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    for prob, f, a, b in zip(probs, plain, alone, mixed):
        for o in (a, b):
            assert o.fp64_fallback
            assert o.termination_reason == TR.TERMINATION_REASON_OPTIMAL
            assert o.refinement_rounds >= 1
            assert o.iteration_count > f.iteration_count
            assert np.array_equal(o.primal_solution, f.primal_solution)
            assert np.array_equal(o.dual_solution, f.dual_solution)
        check_kkt(prob, a, 1e-8)


def test_batch_load_then_run_matches_batch_solve(gpu_ctx):
    """pdlp_batch_load + pdlp_batch_run return bitwise what pdlp_batch_solve
    returns, in FP64 and in FP32 + refinement with correction solves capped
    at 64 iterations (the stalled instances' FP64 solves then start from the
    FP32 instances' device copies); a loaded batch is solved once."""
    say("case: 3 LPs, load + run vs batch_solve, FP64 and FP32 + refinement (inner cap 64)")
    probs = [config_lp(0)[0], transportation_lp(40, 50, seed=6), lp_cases.tiny_two_var()[0]]
    p64 = params(1e-8)
    p64.precision = Precision.FP64
    for p in (p64, params(1e-8, refine_inner_iteration_limit=64)):
        ref = gpu_ctx.batch_solve(probs, p)
        b = gpu_ctx.batch_load(probs, p)
        outs = b.run()
        with pytest.raises(PdlpError):
            b.run()
        b.free()
        for o, r in zip(outs, ref):
            assert o.termination_reason == r.termination_reason
            assert o.iteration_count == r.iteration_count
            assert o.fp64_fallback == r.fp64_fallback
            assert np.array_equal(o.primal_solution, r.primal_solution)
            assert np.array_equal(o.dual_solution, r.dual_solution)
        if p.precision == Precision.FP32_REFINE:
            assert any(o.fp64_fallback for o in outs)


def test_batch_groups_bitwise(gpu_ctx):
    """A batch split into size groups, each on its own streams and host
    thread, returns bitwise what one group returns, in FP64 and in FP32 +
    refinement with correction solves capped at 64 iterations (FP64
    fallbacks in every group)."""
    say("case: 6 LPs, batch_groups 3 vs 1, FP64 and FP32 + refinement (inner cap 64)")
    probs = [config_lp(0)[0], transportation_lp(40, 50, seed=6), lp_cases.tiny_two_var()[0],
             random_lp(n=3000, m=1500, seed=3)[0], random_lp(n=800, m=400, seed=4)[0],
             lp_cases.long_col_lp()[0]]
    p64 = params(1e-8)
    p64.precision = Precision.FP64
    for p in (p64, params(1e-8, refine_inner_iteration_limit=64)):
        p.batch_groups = 1
        ref = gpu_ctx.batch_solve(probs, p)
        p.batch_groups = 3
        out = gpu_ctx.batch_solve(probs, p)
        for o, r in zip(out, ref):
            assert o.termination_reason == r.termination_reason
            assert o.iteration_count == r.iteration_count
            assert o.total_attempts == r.total_attempts
            assert o.fp64_fallback == r.fp64_fallback
            assert np.array_equal(o.primal_solution, r.primal_solution)
            assert np.array_equal(o.dual_solution, r.dual_solution)


def test_early_fallback_bitwise(gpu_ctx):
    """FP64 solves started early (refine_early_fallback: 5 iterations into a
    correction solve capped at 256) return bitwise what starting them at the
    stall returns, and are dropped for instances that refine after all; also
    under an iteration limit of 400 with a cap of 64, where an early solve
    that has run past what is left of the limit is replaced by a fresh one."""
    say("case: 6 LPs, FP32 + refinement, early FP64 fallback vs none (caps 256 and 64)")
    probs = [config_lp(0)[0], transportation_lp(40, 50, seed=6), lp_cases.tiny_two_var()[0],
             random_lp(n=3000, m=1500, seed=3)[0], random_lp(n=800, m=400, seed=4)[0],
             lp_cases.long_col_lp()[0]]
    cases = [(params(1e-8, refine_inner_iteration_limit=256), None),
             (params(1e-8, refine_inner_iteration_limit=64, refine_max_outer=6), 400)]
    for p, limit in cases:
        if limit:
            p.termination_criteria.iteration_limit = limit
        p.refine_early_fallback = 0
        ref = gpu_ctx.batch_solve(probs, p)
        p.refine_early_fallback = 5
        out = gpu_ctx.batch_solve(probs, p)
        say(f"  cap {p.refine_inner_iteration_limit}: fallbacks {[o.fp64_fallback for o in out]}, "
            f"rounds {[o.refinement_rounds for o in out]}")
        for o, r in zip(out, ref):
            assert o.termination_reason == r.termination_reason
            assert o.iteration_count == r.iteration_count
            assert o.total_attempts == r.total_attempts
            assert o.refinement_rounds == r.refinement_rounds
            assert o.fp64_fallback == r.fp64_fallback
            assert np.array_equal(o.primal_solution, r.primal_solution)
            assert np.array_equal(o.dual_solution, r.dual_solution)
        assert any(o.fp64_fallback for o in out)


def test_refined_small_cases(gpu_ctx):
    say("case: tiny LPs, FP32 + refinement")
    prob, opt = lp_cases.tiny_two_var()
    out = gpu_ctx.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    np.testing.assert_allclose(out.primal_solution, [2.0, 0.0], atol=1e-7)
    out = gpu_ctx.solve(lp_cases.primal_infeasible(), params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_PRIMAL_INFEASIBLE
    prob, opt = lp_cases.long_row_lp(6000)
    out = gpu_ctx.solve(prob, params(1e-8))
    assert out.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    check_kkt(prob, out, 1e-8)


def test_refined_batch_alone_vs_mixed(gpu_ctx):
    say("case: FP32 + refinement, batch of 4 vs alone (bitwise)")
    probs = [random_lp(n=n, m=n // 2, seed=s)[0] for n, s in ((300, 21), (2000, 22), (9000, 23), (700, 24))]
    p = params(1e-8)
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    for prob, a, b in zip(probs, alone, mixed):
        assert a.termination_reason == b.termination_reason == TR.TERMINATION_REASON_OPTIMAL
        assert a.iteration_count == b.iteration_count
        assert a.refinement_rounds == b.refinement_rounds
        assert np.array_equal(a.primal_solution, b.primal_solution)
        assert np.array_equal(a.dual_solution, b.dual_solution)
        check_kkt(prob, b, 1e-8)


def test_capped_round_continues_when_it_made_progress(gpu_ctx):
    """refine_continue_progress = 10 on two LPs of the 512-LP batch (seeds
    41 and 510) whose first correction solve stops at its 10,000-iteration
    cap after cutting the refined point's error more than ten times: the
    refinement goes on from that point instead of falling back at once.
    Results are the same alone and in a batch (bitwise) and meet the
    tolerance; more rounds run than with the fallback at the first capped
    round, and seed 510 then refines to optimal without an FP64 solve."""
    say("case: batch LPs 41 and 510, FP32 + refinement, continue when a capped round cut the error 10x")
    sizes = batch_sizes(512)
    probs = [random_lp(n=int(sizes[s]), m=int(sizes[s]) // 2, seed=s)[0] for s in (41, 510)]
    p = params(1e-8, refine_continue_progress=10)
    p0 = params(1e-8)
    alone = [gpu_ctx.solve(q, p) for q in probs]
    mixed = gpu_ctx.batch_solve(probs, p)
    stop_at_once = [gpu_ctx.solve(q, p0) for q in probs]
    say(f"  rounds {[o.refinement_rounds for o in alone]} (stall at once: "
        f"{[o.refinement_rounds for o in stop_at_once]}), fallbacks "
        f"{[o.fp64_fallback for o in alone]}")
    for prob, a, b in zip(probs, alone, mixed):
        assert a.termination_reason == b.termination_reason == TR.TERMINATION_REASON_OPTIMAL
        assert a.iteration_count == b.iteration_count
        assert a.refinement_rounds == b.refinement_rounds
        assert a.fp64_fallback == b.fp64_fallback
        assert np.array_equal(a.primal_solution, b.primal_solution)
        assert np.array_equal(a.dual_solution, b.dual_solution)
        check_kkt(prob, a, 1e-8)
    assert all(o.fp64_fallback for o in stop_at_once)
    assert all(a.refinement_rounds > s.refinement_rounds for a, s in zip(alone, stop_at_once))
    assert not alone[1].fp64_fallback
