"""The step kernels' tile parameters and pdlp_tune.

* slice_unroll (4, 8, 16 or each kernel's default), slice_waves and the
  occupancy caps change only how the work is scheduled: the results are
  bitwise those of the defaults.
* lane_budget changes the slicing, hence the order of each row's
  compensated sum: the results are still bitwise equal to the CPU oracle's
  on these LPs (the bar of tests/test_gpu_parity.py).
* pdlp_tune returns one of its candidates, caches it per LP shape, and the
  chosen parameters solve a batch to optimality with each result bitwise
  that of a lone solve with the same parameters.
"""
import sys

import numpy as np
import pytest

import oracle as O
from cupdlp import (PdhgParameters, Precision, TerminationReason as TR,
                    construct_termination_criteria)
from cupdlp.generators import config_lp, random_lp, transportation_lp
from test_gpu_parity import compare_runs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params(tol=1e-12, iteration_limit=2**31 - 1, **kw):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit)
    return PdhgParameters(termination_criteria=tc, **kw)


def test_unroll_and_waves_change_no_result(gpu_ctx):
    say("case: slice_unroll 4/8/16 x slice_waves 1/2/4 x occupancy caps against the defaults "
        "(unroll 8, 16 for the FP64 dual kernel), n=2000 LP and scaled Zipf LP, FP64 and FP32")
    for prob in (config_lp(0)[0], config_lp(1, scale=0.02)[0]):
        for prec in (Precision.FP64, Precision.FP32_REFINE):
            base = gpu_ctx.solve(prob, params(1e-8, iteration_limit=600, precision=prec))
            for u, w, po, do in ((4, 4, 0, 0), (8, 2, 0, 0), (4, 1, 0, 0), (8, 1, 0, 0),
                                 (8, 4, 8, 8), (8, 4, 3, 6), (8, 2, 4, 2), (4, 1, 1, 16),
                                 (8, 4, 0, 0), (16, 4, 0, 0), (16, 1, 5, 3)):
                g = gpu_ctx.solve(prob, params(1e-8, iteration_limit=600, precision=prec,
                                               slice_unroll=u, slice_waves=w,
                                               primal_occupancy=po, dual_occupancy=do))
                compare_runs(g, base)


def test_tile_parameters_out_of_range_are_refused(gpu_ctx):
    say("case: tile parameters out of range refused (code 2) before any device work; "
        "slice_unroll 16 accepted")
    from cupdlp.solver import PdlpError
    prob = config_lp(0, scale=0.1)[0]
    for kw in ({"slice_unroll": 12}, {"slice_unroll": 2}, {"slice_waves": 3},
               {"lane_budget": 8}, {"lane_budget": 2048}, {"primal_occupancy": 17},
               {"dual_occupancy": -1}):
        with pytest.raises(PdlpError) as ei:
            gpu_ctx.solve(prob, params(1e-6, iteration_limit=10, **kw))
        assert ei.value.code == 2, (kw, ei.value.code)
    ok = gpu_ctx.solve(prob, params(1e-12, iteration_limit=10, slice_unroll=16))
    assert ok.termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    assert ok.iteration_count == 10


@pytest.mark.parametrize("lane_budget", [16, 32, 128])
def test_lane_budget_matches_the_oracle(gpu_ctx, lane_budget):
    say(f"case: lane_budget {lane_budget} vs the oracle (n=2000 LP to 1e-8, 60x60 "
        "transportation LP to 1e-6, scaled Zipf LP 400 iterations)")
    cases = ((config_lp(0)[0], params(1e-8)),
             (transportation_lp(60, 60, seed=5), params(1e-6)),
             (config_lp(1, scale=0.02)[0], params(iteration_limit=400)))
    for prob, p in cases:
        p.lane_budget = lane_budget
        compare_runs(gpu_ctx.solve(prob, p), O.solve(prob, p))


def test_tune_picks_a_candidate_and_caches_it(gpu_ctx):
    say("case: pdlp_tune on a 24-LP batch (FP64), then its choice solves the batch")
    probs = [random_lp(n=n, m=n // 2, seed=60 + k)[0]
             for k, n in enumerate([500, 900, 2000, 4000, 8000, 16000] * 4)]
    p = params(1e-6)
    t = gpu_ctx.tune(probs, p, trial_iterations=50)
    # FP64: one more unroll trial (8 for the dual kernel too) unless 4 was taken
    assert t["candidates"] == (10 if t["slice_unroll"] == 4 else 11)
    assert t["lane_budget"] in (32, 64, 128) and t["slice_unroll"] in (0, 4, 8)
    assert t["slice_waves"] in (2, 4)
    assert t["primal_occupancy"] in (0, 3, 4, 5, 6, 8) and t["dual_occupancy"] in (0, 3, 4, 5, 6, 8)
    assert 0 < t["kernel_us_per_iteration"] <= t["default_us_per_iteration"]
    again = gpu_ctx.tune(probs, p, trial_iterations=50)
    assert again["candidates"] == 0
    keys = ("lane_budget", "slice_unroll", "slice_waves", "primal_occupancy", "dual_occupancy")
    assert all(again[k] == t[k] for k in keys)
    for k in keys:
        setattr(p, k, t[k])
    outs = gpu_ctx.batch_solve(probs, p)
    for q, b in list(zip(probs, outs))[::5]:
        a = gpu_ctx.solve(q, p)
        assert b.termination_reason == TR.TERMINATION_REASON_OPTIMAL
        assert a.iteration_count == b.iteration_count
        assert np.array_equal(a.primal_solution, b.primal_solution)
