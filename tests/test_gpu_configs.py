"""GPU tests at the benchmark configurations' own sizes.

* The n = 1,000,000 Zipf-degree LP (m = 500,000, ~12.5M nonzeros, row and
  column scales over 6 and 4 decades): its first 96 iterations in FP64,
  bitwise against the CPU oracle (iterates, step sizes, primal weights,
  restart choices, every KKT statistic); and its generator at n = 20,000
  solved to 1e-4, bitwise the oracle's whole 474k-iteration solve.
* The 512-LP batch (seeds 0..511, n log-uniform in [500, 200,000]) solved
  batched in FP32 + FP64 refinement to its tolerance 1e-6: every instance
  OPTIMAL; the reference's optimality test (src/termination.jl:142-170, L2,
  eps_abs = eps_rel = 1e-6) recomputed with numpy on every returned point;
  the objective against the generator's certified optimum; the oracle's
  KKT statistics on a sample; eight instances bitwise equal to lone solves.
* A served stream after the served configuration's mix: random LPs and 10 %
  transportation LPs, tolerances 1e-4 (70 %) and 1e-8 (30 %), staggered
  arrivals; each result bitwise equal to a lone solve, in FP64 and in
  FP32 + refinement.
"""
import sys

import numpy as np
import pytest

import oracle as O
from cupdlp import (PdhgParameters, Precision, Server, TerminationReason as TR,
                    construct_termination_criteria)
from cupdlp.generators import batch_sizes, config_lp, random_lp, transportation_lp
from cupdlp.kkt import relative_kkt
from test_gpu_parity import compare_runs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params(tol, iteration_limit=2**31 - 1, precision=Precision.FP64, **kw):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit)
    return PdhgParameters(termination_criteria=tc, precision=precision, **kw)


def rel_kkt(prob, x, y):
    """The reference's relative KKT measure of (x, y) (cupdlp.kkt, numpy)."""
    return relative_kkt(prob, x, y)


def test_zipf_lp_n20000_to_tolerance(gpu_ctx):
    """The n = 1,000,000 LP's generator at n = 20,000, solved to 1e-4 in
    FP64: OPTIMAL after exactly the oracle's 473,856-odd iterations with
    bitwise the oracle's final objectives (tests/golden/zipf_scaled_oracle.json,
    a 5-minute oracle run), and within the tolerance of HiGHS's optimum."""
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "zipf_scaled_oracle.json")) as f:
        ref = json.load(f)
    with open(os.path.join(here, "golden", "highs_objectives.json")) as f:
        h = json.load(f)["objectives"]["zipf_scaled_n20000"]["objective"]
    say(f"case: Zipf-degree LP n=20,000 to {ref['tolerance']} in FP64 "
        f"(the oracle: {ref['iteration_count']} iterations)")
    prob, _ = config_lp(1, scale=0.02)
    g = gpu_ctx.solve(prob, params(ref["tolerance"]))
    ci = g.iteration_stats[-1].convergence_information[0]
    say(f"  GPU: {g.termination_reason.name}, {g.iteration_count} iterations, "
        f"{g.solve_time_sec:.1f} s")
    assert int(g.termination_reason) == ref["termination_reason"]
    assert g.termination_reason == TR.TERMINATION_REASON_OPTIMAL
    assert g.iteration_count == ref["iteration_count"]
    assert ci.primal_objective == ref["primal_objective"]
    assert ci.dual_objective == ref["dual_objective"]
    assert abs(ci.primal_objective - h) <= ref["tolerance"] * (1 + abs(h))
    assert rel_kkt(prob, g.primal_solution, g.dual_solution) <= ref["tolerance"]


def test_c1_full_size_first_96_iterations(gpu_ctx):
    say("case: n=1,000,000 Zipf-degree scaled LP, 96 iterations, FP64 vs the oracle")
    prob, _ = config_lp(1)
    assert prob.num_variables == 1_000_000 and prob.constraint_matrix.nnz > 12_000_000
    p = params(1e-12, iteration_limit=96)
    g = gpu_ctx.solve(prob, p, want_current=True)
    say("  GPU done, running the oracle")
    o = O.solve(prob, p, want_current=True)
    assert g.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    compare_runs(g, o)
    assert np.array_equal(g.primal_current, o.primal_current)
    assert np.array_equal(g.dual_current, o.dual_current)


def test_c3_full_batch_to_tolerance(gpu_ctx):
    tol = 1e-6
    say("case: the 512-LP batch, FP32 + refinement, batched, to 1e-6")
    sizes = batch_sizes(512)
    gens = [random_lp(n=int(n), m=int(n) // 2, seed=i) for i, n in enumerate(sizes)]
    probs = [g[0] for g in gens]
    p = params(tol, precision=Precision.FP32_REFINE, record_iteration_stats=False)
    outs = gpu_ctx.batch_solve(probs, p)
    say(f"  solved; {sum(o.fp64_fallback for o in outs)} through the FP64 fallback")
    worst_kkt, worst_obj = 0.0, 0.0
    for i, ((prob, cert), o) in enumerate(zip(gens, outs)):
        assert o.termination_reason == TR.TERMINATION_REASON_OPTIMAL, i
        k = rel_kkt(prob, o.primal_solution, o.dual_solution)
        assert k < tol * (1 + 1e-6), (i, k)
        obj = prob.objective_vector @ o.primal_solution
        e = abs(obj - cert.objective) / (1 + abs(cert.objective))
        assert e <= 50 * tol, (i, e)
        worst_kkt, worst_obj = max(worst_kkt, k), max(worst_obj, e)
    say(f"  max relative KKT {worst_kkt:.3g}, max objective error vs certificate {worst_obj:.3g}")
    # the oracle's statistics agree with numpy's on a sample
    for i in range(0, 512, 32):
        ci, _ = O.point_stats(probs[i], outs[i].primal_solution, outs[i].dual_solution, tol, tol)
        k = rel_kkt(probs[i], outs[i].primal_solution, outs[i].dual_solution)
        kk = max(ci.relative_l2_primal_residual, ci.relative_l2_dual_residual,
                 ci.relative_optimality_gap)
        assert abs(kk - k) <= 1e-6 * k + 1e-15, (i, kk, k)
    # eight instances alone give the same result bitwise
    for i in (0, 3, 77, 128, 205, 300, 421, 511):
        a = gpu_ctx.solve(probs[i], p)
        b = outs[i]
        assert a.termination_reason == b.termination_reason, i
        assert a.iteration_count == b.iteration_count, i
        assert a.fp64_fallback == b.fp64_fallback, i
        assert np.array_equal(a.primal_solution, b.primal_solution), i
        assert np.array_equal(a.dual_solution, b.dual_solution), i


def served_mix(count=20, seed=4):
    """The served configuration's mix at test size: 90 % random LPs (n
    log-uniform in [1,000, 20,000]), 10 % transportation LPs (30..80 per
    side), tolerance 1e-4 (70 %) or 1e-8 (30 %)."""
    rng = np.random.default_rng(seed)
    probs, tols = [], []
    for k in range(count):
        s = int(rng.integers(0, 2**31 - 1))
        if k % 10 == 3:
            a, b = (int(v) for v in rng.integers(30, 81, size=2))
            probs.append(transportation_lp(a, b, seed=s))
        else:
            n = int(np.exp(rng.uniform(np.log(1000), np.log(20000))))
            probs.append(random_lp(n=n, m=n // 2, seed=s)[0])
        tols.append(1e-4 if rng.random() < 0.7 else 1e-8)
    return probs, tols


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_REFINE])
def test_c4_style_stream_alone_vs_mixed(gpu_ctx, precision):
    say(f"case: served stream of 20 requests (2 transportation), tolerances 1e-4/1e-8, {precision.name}")
    probs, tols = served_mix()
    assert any(t == 1e-8 for t in tols) and any(t == 1e-4 for t in tols)
    alone = [gpu_ctx.solve(q, params(t, precision=precision)) for q, t in zip(probs, tols)]
    srv = Server(gpu_ctx, params(1e-4, precision=precision), hbm_budget_bytes=200e9)
    out, ids = {}, []
    try:
        for lo, hi in ((0, 5), (5, 12), (12, 20)):
            ids += [srv.submit(probs[i], tols[i]) for i in range(lo, hi)]
            for _ in range(3):
                for f in srv.poll():
                    out[f] = srv.result(f)
        for _ in range(2_000_000):
            if len(out) == len(ids):
                break
            for f in srv.poll():
                out[f] = srv.result(f)
    finally:
        srv.close()
    assert len(out) == len(ids)
    for k, rid in enumerate(ids):
        a, b = alone[k], out[rid]
        assert b.termination_reason == TR.TERMINATION_REASON_OPTIMAL, k
        assert a.termination_reason == b.termination_reason, k
        assert a.iteration_count == b.iteration_count, k
        assert a.total_attempts == b.total_attempts, k
        assert np.array_equal(a.primal_solution, b.primal_solution), k
        assert np.array_equal(a.dual_solution, b.dual_solution), k
        assert rel_kkt(probs[k], b.primal_solution, b.dual_solution) < tols[k] * (1 + 1e-6), k
