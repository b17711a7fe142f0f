"""GPU tests of the continuous-batching request loop (pdlp_serve_*).

The bar: a request's result is the same whether it runs alone or arrives
while other requests are on the device -- bitwise (x, y, iteration and
attempt counts, termination reason, final statistics), since instances
packed into the same launches share no arithmetic.  A lone solve with
eps_optimal_absolute = eps_optimal_relative = the request's tolerance is
the reference for "alone".
"""
import numpy as np
import pytest

import lp_cases
from cupdlp import PdhgParameters, PdlpError, Precision, Server, construct_termination_criteria
from cupdlp.generators import random_lp, transportation_lp

pytestmark = pytest.mark.gpu


def params(tol, precision=Precision.FP64):
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol)
    return PdhgParameters(termination_criteria=tc, precision=precision)


def same_result(a, b):
    assert a.termination_reason == b.termination_reason
    assert a.iteration_count == b.iteration_count
    assert a.total_attempts == b.total_attempts
    assert np.array_equal(a.primal_solution, b.primal_solution)
    assert np.array_equal(a.dual_solution, b.dual_solution)
    ca = a.iteration_stats[-1].convergence_information[0]
    cb = b.iteration_stats[-1].convergence_information[0]
    assert ca.primal_objective == cb.primal_objective
    assert ca.l2_primal_residual == cb.l2_primal_residual
    assert ca.l2_dual_residual == cb.l2_dual_residual


def stream():
    sizes = [300, 4000, 900, 12000, 2500, 600, 7000, 1500]
    probs = [random_lp(n=n, m=n // 2, seed=40 + i)[0] for i, n in enumerate(sizes)]
    probs.append(transportation_lp(30, 40, seed=3))
    tols = [1e-4 if i % 3 else 1e-8 for i in range(len(probs))]
    return probs, tols


def drain(srv, ids, out, max_rounds=1_000_000):
    for _ in range(max_rounds):
        if all(i in out for i in ids):
            return
        for f in srv.poll():
            out[f] = srv.result(f)
    raise AssertionError("requests did not finish")


def test_serve_alone_vs_mixed(gpu_ctx):
    print("case: 9 requests (random LPs n=300..12000, one transportation LP), tolerances 1e-4/1e-8, "
          "staggered arrivals", flush=True)
    probs, tols = stream()
    alone = [gpu_ctx.solve(p, params(t)) for p, t in zip(probs, tols)]
    srv = Server(gpu_ctx, params(1e-4))
    try:
        out = {}
        ids = [srv.submit(probs[i], tols[i]) for i in range(3)]
        for _ in range(2):
            for f in srv.poll():
                out[f] = srv.result(f)
        ids += [srv.submit(probs[i], tols[i]) for i in range(3, 6)]
        while not out:
            for f in srv.poll():
                out[f] = srv.result(f)
        ids += [srv.submit(probs[i], tols[i]) for i in range(6, len(probs))]
        drain(srv, ids, out)
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        assert out[rid].termination_string == "OPTIMAL"
        same_result(out[rid], alone[k])
        assert out[rid].solve_time_sec > 0.0


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_REFINE])
def test_serve_degenerate_shapes_alone_vs_mixed(gpu_ctx, precision):
    print(f"case: served stream of degenerate / structured LPs vs alone, {precision.name} "
          "(empty rows and column, fixed and one-sided bounds, no equalities, long column "
          "and row, m = 0), tolerances 1e-4 / 1e-8", flush=True)
    probs = [lp_cases.empty_row_and_column_lp()[0], lp_cases.fixed_variables_lp()[0],
             lp_cases.one_sided_bounds_lp()[0], random_lp(n=300, m=150, num_eq=0, seed=41)[0],
             lp_cases.long_col_lp(3000)[0], lp_cases.long_row_lp(6000)[0],
             lp_cases.no_rows()[0]]
    tols = [1e-8 if i % 2 else 1e-4 for i in range(len(probs))]
    alone = [gpu_ctx.solve(p, params(t, precision)) for p, t in zip(probs, tols)]
    srv = Server(gpu_ctx, params(1e-4, precision))
    try:
        out = {}
        ids = [srv.submit(probs[i], tols[i]) for i in range(4)]
        for _ in range(3):
            for f in srv.poll():
                out[f] = srv.result(f)
        ids += [srv.submit(probs[i], tols[i]) for i in range(4, len(probs))]
        drain(srv, ids, out)
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        assert out[rid].termination_string == "OPTIMAL", k
        same_result(out[rid], alone[k])


def test_serve_one_slot_and_budget(gpu_ctx):
    print("case: one slot at a time (admission waits for retirement); a request over the HBM budget",
          flush=True)
    probs, tols = stream()
    probs, tols = probs[:4], tols[:4]
    alone = [gpu_ctx.solve(p, params(t)) for p, t in zip(probs, tols)]
    srv = Server(gpu_ctx, params(1e-4), max_slots=1)
    try:
        out = {}
        ids = [srv.submit(p, t) for p, t in zip(probs, tols)]
        drain(srv, ids, out)
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        same_result(out[rid], alone[k])
    small = Server(gpu_ctx, params(1e-4), hbm_budget_bytes=4096.0)
    try:
        with pytest.raises(PdlpError) as e:
            small.submit(probs[1], 1e-4)
        assert e.value.code == 5
    finally:
        small.close()


def test_serve_fp32_refine_alone_vs_mixed(gpu_ctx):
    print("case: FP32 + refinement server, 3 requests to 1e-8", flush=True)
    probs = [random_lp(n=n, m=n // 2, seed=60 + i)[0] for i, n in enumerate((800, 3000, 1500))]
    alone = [gpu_ctx.solve(p, params(1e-8, Precision.FP32_REFINE)) for p in probs]
    srv = Server(gpu_ctx, params(1e-4, Precision.FP32_REFINE))
    try:
        out = {}
        ids = [srv.submit(probs[0], 1e-8)]
        for _ in range(3):
            for f in srv.poll():
                out[f] = srv.result(f)
        ids += [srv.submit(p, 1e-8) for p in probs[1:]]
        drain(srv, ids, out)
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        assert out[rid].termination_string == "OPTIMAL"
        same_result(out[rid], alone[k])


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_REFINE])
def test_serve_size_classes_bitwise(gpu_ctx, precision):
    """Requests of three size classes (nonzeros below 2^17, below 2^20 and
    above) run in engines of their own, side by side: each result is bitwise
    the lone solve's and the same as with a single engine (batch_groups 1)."""
    print(f"case: served requests n = 1,000..140,000 in 3 size classes vs 1, {precision.name}",
          flush=True)
    sizes = [1000, 140_000, 3000, 20_000, 40_000, 1500]
    probs = [random_lp(n=n, m=n // 2, seed=80 + i)[0] for i, n in enumerate(sizes)]
    tols = [1e-6, 1e-4, 1e-8, 1e-6, 1e-4, 1e-8]
    alone = [gpu_ctx.solve(p, params(t, precision)) for p, t in zip(probs, tols)]
    res = {}
    for classes in (3, 1):
        sp = params(1e-4, precision)
        sp.batch_groups = classes
        srv = Server(gpu_ctx, sp)
        try:
            out = {}
            ids = [srv.submit(probs[i], tols[i]) for i in range(3)]
            for _ in range(3):
                for f in srv.poll():
                    out[f] = srv.result(f)
            ids += [srv.submit(probs[i], tols[i]) for i in range(3, len(probs))]
            drain(srv, ids, out)
        finally:
            srv.close()
        res[classes] = [out[i] for i in ids]
    for k in range(len(probs)):
        same_result(res[3][k], alone[k])
        same_result(res[1][k], alone[k])


def test_serve_fp32_stalled_falls_back_to_fp64(gpu_ctx):
    print("case: FP32 + refinement server with inner cap 64: stalled requests finish in FP64, "
          "bitwise the plain FP64 solve", flush=True)
    from cupdlp.generators import config_lp
    probs = [config_lp(0)[0], random_lp(n=900, m=450, seed=70)[0], transportation_lp(30, 40, seed=3)]
    plain = [gpu_ctx.solve(p, params(1e-8)) for p in probs]
    sp = params(1e-4, Precision.FP32_REFINE)
    sp.refine_inner_iteration_limit = 64
    srv = Server(gpu_ctx, sp)
    try:
        out = {}
        ids = [srv.submit(probs[0], 1e-8)]
        for _ in range(4):
            for f in srv.poll():
                out[f] = srv.result(f)
        ids += [srv.submit(p, 1e-8) for p in probs[1:]]
        drain(srv, ids, out)
    finally:
        srv.close()
    for k, rid in enumerate(ids):
        o = out[rid]
        assert o.termination_string == "OPTIMAL"
        assert o.fp64_fallback
        assert o.iteration_count > plain[k].iteration_count
        assert np.array_equal(o.primal_solution, plain[k].primal_solution)
        assert np.array_equal(o.dual_solution, plain[k].dual_solution)


def test_serve_request_limits(gpu_ctx):
    """Per-request limits (pdlp_serve_submit_limits): an iteration limit gives
    the lone solve with that limit, bitwise; a deadline that passes while
    queued or running ends the request with TIME_LIMIT; the others are
    unaffected."""
    print("case: per-request iteration and time limits", flush=True)
    from cupdlp import TerminationReason as TR
    probs, tols = stream()
    tc = construct_termination_criteria(eps_optimal_absolute=1e-8, eps_optimal_relative=1e-8,
                                        iteration_limit=50)
    lim50 = gpu_ctx.solve(probs[3], PdhgParameters(termination_criteria=tc))
    alone0 = gpu_ctx.solve(probs[0], params(tols[0]))
    big, _ = random_lp(n=200_000, m=100_000, seed=9)
    srv = Server(gpu_ctx, params(1e-4))
    try:
        out = {}
        a = srv.submit(probs[0], tols[0])
        b = srv.submit(probs[3], 1e-8, iteration_limit=50)
        c = srv.submit(probs[1], 1e-8, time_sec_limit=1e-9)  # expires before admission
        d = srv.submit(big, 1e-12, time_sec_limit=0.3)       # expires while running
        drain(srv, [a, b, c, d], out)
    finally:
        srv.close()
    same_result(out[a], alone0)
    assert out[b].termination_reason == TR.TERMINATION_REASON_ITERATION_LIMIT
    assert out[b].iteration_count == 50
    same_result(out[b], lim50)
    assert out[c].termination_reason == TR.TERMINATION_REASON_TIME_LIMIT
    assert out[c].iteration_count == 0 and not out[c].primal_solution.any()
    assert out[d].termination_reason == TR.TERMINATION_REASON_TIME_LIMIT
    assert out[d].solve_time_sec >= 0.3 and out[d].iteration_count > 0


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_REFINE])
def test_serve_cancel(gpu_ctx, precision):
    """pdlp_serve_cancel of a running and of a queued request: both finish
    with reason OTHER and no solution; a request running beside them and
    the requests after them keep bitwise the results of lone solves."""
    print(f"case: cancel a running and a queued request, {precision.name}", flush=True)
    from cupdlp import TerminationReason as TR
    probs, tols = stream()
    keep = {3: 1e-8, 0: tols[0], 5: tols[5]}  # request 3 (n = 12,000) at 1e-8 runs long
    alone = {k: gpu_ctx.solve(probs[k], params(t, precision)) for k, t in keep.items()}
    big, _ = random_lp(n=150_000, m=75_000, seed=10)
    srv = Server(gpu_ctx, params(1e-4, precision), max_slots=2)
    try:
        out = {}
        run = srv.submit(big, 1e-12)
        ids = {3: srv.submit(probs[3], 1e-8)}
        queued = srv.submit(big, 1e-12)  # both slots taken: it waits
        for _ in range(2):
            for f in srv.poll():
                out[f] = srv.result(f)
        assert not out  # nothing finished yet
        srv.cancel(run)
        srv.cancel(queued)
        with pytest.raises(PdlpError):
            srv.cancel(10_000)
        ids[0] = srv.submit(probs[0], keep[0])
        ids[5] = srv.submit(probs[5], keep[5])
        drain(srv, list(ids.values()) + [run, queued], out)
        with pytest.raises(PdlpError):
            srv.cancel(run)  # already finished
    finally:
        srv.close()
    for k, rid in ids.items():
        same_result(out[rid], alone[k])
    for rid in (run, queued):
        assert out[rid].termination_reason == TR.TERMINATION_REASON_OTHER
        assert not out[rid].primal_solution.any()
    assert out[run].iteration_count > 0 and out[queued].iteration_count == 0
