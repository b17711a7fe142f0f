#!/usr/bin/env python3
"""Benchmark of the MI355X restarted-PDHG engine (one GPU).

Default workload (the batched configuration of the headline metric): 512 LPs
from the random-LP generator (seeds 0..511, n log-uniform in [500, 200,000],
m = n/2, 8 nonzeros per column), solved together by the batched engine in
FP32 with FP64 iterative refinement to 1e-8 relative KKT (the reference's
own optimality test, eps_abs = eps_rel; refinements that stall are finished
in FP64).  A "step" is one solve of the whole batch; `value` is solves/s
over the K timed steps.  Each step first loads the LPs onto the device
(pdlp_batch_load: host layout and the copy to HBM, untimed, as optimize()
copies the problem before its solve clock starts,
src/primal_dual_hybrid_gradient_gpu.jl:488-490 and :602); the timed region
is pdlp_batch_run -- rescaling on the device, the solves, the FP64 solves of
stalled refinements and the result download.  The rate with the load
included is reported beside it.

Reported beside it:
  * plain FP64 one-at-a-time solves/s on the same GPU (the whole batch, or a
    prefix if the budget runs out), the speed-up over it, and its p50/p99
    latency (completion times with the batch submitted at once and solved
    in order) and p99 per-solve time;
  * the same batch in plain FP64, batched (one timed solve);
  * p50/p99 latency of the instances within a batch step (time from the
    start of the run to each result);
  * per-result error of every returned point: the reference's relative KKT
    measure (max, median) and the objective error against the generator's
    certified optimum;
  * `roofline`: the dominant iteration kernel's algorithmic bytes per launch
    over its HIP-event duration (events on the library's own stream);
    `fallback_kernels`: the same for the FP64 engines that finish stalled
    refinements; `aggregate_step_kernels`: the algorithmic bytes of every
    step-kernel launch of the timed runs over their wall time;
  * `served`: a bounded served stream in the same precision (request loop,
    Poisson arrivals), solves/s and p50/p99 latency against plain FP64 one
    at a time on the same requests;
  * `cpu_baseline`: the CPU oracle (same restarted PDHG, OpenMP) on a bounded
    sample of the same LPs on this host's cores.

Other workloads: --workload serve (a Poisson request stream through the
continuous-batching loop, p50/p99 latency), --workload c1 (n = 1,000,000 LP, fixed iteration count, for
the FP64 SpMV roofline), --workload c2 (2,000 x 2,000 transportation LP),
--workload c0 (n = 2,000 LP), --workload copy (a plain 1 GiB device copy, the
calibration kernel for the rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cupdlp.jl_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "solves/s + p99 latency to 1e-8 rel KKT, FP32+refine batched vs plain FP64"

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (2 * FETCH + WRITE, factors from a plain 1 GiB copy on the same box), see
# profiles/hbm_traffic_counters.md and profiles/batch_kernels_512lp_gbs.md
# (the batch rows).  Keyed by (workload, precision, kernel)
# and used only when this run's algorithmic bytes per launch are within
# 0.5 % of the profiled run's (the same workload: when a batch instance
# leaves the launches depends on the host's round timing, so the mix moves
# by a few hundredths of a percent between runs); the profiled ratio of
# traffic to algorithmic bytes is then applied to this run's bytes.
MEASURED_TRAFFIC = {
    ("batch", "fp64", "k_dual"): (69376582.0, 96.262e6),
    ("batch", "fp64", "k_primal"): (85985015.0, 104.654e6),
    ("batch", "fp32", "k_dual"): (172089502.0, 250.583e6),
    ("batch", "fp32", "k_primal"): (207292512.0, 264.742e6),
    ("c1", "fp64", "k_dual"): (195460496.0, 286.819e6),
    ("c1", "fp64", "k_primal"): (229460496.0, 247.593e6),
    ("c1", "fp32", "k_dual"): (123640332.0, 175.487e6),
    ("c2", "fp64", "k_primal"): (400032004.0, 423.452e6),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


BATCH_GROUPS = 0  # --batch-groups
REFINE_INNER_TOL = 0.0  # --refine-inner-tol (0: the library default)
REFINE_EARLY = 0  # --refine-early-fallback (0: the library default)
REFINE_WARM = -1  # --refine-warm-start (-1: the library default)
REFINE_CAP = 0  # --refine-cap (0: the library default)
REFINE_CONTINUE = -1  # --refine-continue (-1: the library default)
OCCUPANCY = ()  # --slice-occupancy P,D (empty: the tuner's choice or the library default)


def make_params(tol, precision, iteration_limit=2**31 - 1, time_limit=float("inf")):
    from cupdlp import PdhgParameters, Precision, construct_termination_criteria
    tc = construct_termination_criteria(eps_optimal_absolute=tol, eps_optimal_relative=tol,
                                        iteration_limit=iteration_limit, time_sec_limit=time_limit)
    p = PdhgParameters(termination_criteria=tc, record_iteration_stats=False,
                       precision=Precision.FP32_REFINE if precision == "fp32" else Precision.FP64,
                       batch_groups=BATCH_GROUPS)
    if REFINE_INNER_TOL > 0:
        p.refine_inner_tolerance = REFINE_INNER_TOL
    p.refine_early_fallback = REFINE_EARLY
    if REFINE_WARM >= 0:
        p.refine_warm_start = REFINE_WARM
    if REFINE_CAP > 0:
        p.refine_inner_iteration_limit = REFINE_CAP
    if REFINE_CONTINUE >= 0:
        p.refine_continue_progress = REFINE_CONTINUE
    for k, v in TILES.get(precision, {}).items():
        setattr(p, k, v)
    if OCCUPANCY:
        p.primal_occupancy, p.dual_occupancy = OCCUPANCY
    return p


# Step-kernel tile parameters chosen by pdlp_tune, per precision (--tune).
TILES = {}


def tune(ctx, problems, precision, tol, trial_iterations=100):
    """pdlp_tune on these LPs before anything is timed: sets TILES[precision]
    and returns the tuner's report."""
    t = ctx.tune(problems, make_params(tol, precision), trial_iterations)
    TILES[precision] = {k: t[k] for k in ("lane_budget", "slice_unroll", "slice_waves",
                                          "primal_occupancy", "dual_occupancy")}
    log(f"tuned {precision}: {TILES[precision]} ({t['kernel_us_per_iteration']:.1f} us vs "
        f"{t['default_us_per_iteration']:.1f} us default per iteration, {t['candidates']} "
        f"candidates, {t['seconds']:.1f} s)")
    return {"chosen": TILES[precision], "kernel_us_per_iteration": round(t["kernel_us_per_iteration"], 2),
            "default_us_per_iteration": round(t["default_us_per_iteration"], 2),
            "candidates": t["candidates"], "trial_iterations": trial_iterations,
            "seconds": round(t["seconds"], 2)}


def batch_problems(count):
    """The batch LPs and their certified optimal objectives."""
    from cupdlp.generators import batch_sizes, random_lp
    sizes = batch_sizes(count)
    gens = [random_lp(n=int(n), m=int(n) // 2, seed=i) for i, n in enumerate(sizes)]
    return [g[0] for g in gens], [g[1].objective for g in gens]


def accuracy(problems, outs, objectives):
    """Per-result error of the returned points: the reference's relative KKT
    measure (cupdlp.kkt; optimal at eps iff < eps) and the objective error
    against the generator's certified optimum (an FP64 optimal solution by
    construction), relative to 1 + |optimum|."""
    from cupdlp.kkt import kkt_parts
    kkt, err = [], []
    for prob, o, ref in zip(problems, outs, objectives):
        rp, rd, rg, p, _ = kkt_parts(prob, o.primal_solution, o.dual_solution)
        kkt.append(max(rp, rd, rg))
        err.append(abs(p - ref) / (1.0 + abs(ref)))
    return {"max_rel_kkt": float(np.max(kkt)), "median_rel_kkt": float(np.median(kkt)),
            "max_rel_obj_error_vs_certificate": float(np.max(err)),
            "median_rel_obj_error_vs_certificate": float(np.median(err))}


def fallback_kernels(prof):
    """The FP64 engines that finish stalled FP32 + refinement instances:
    launches, average durations and achieved GB/s of their step kernels."""
    if prof.fallback_primal_launches == 0:
        return None
    pm, dm = prof.fallback_primal_ms, prof.fallback_dual_ms
    if pm == 0.0:  # counted, not timed (--profile-fallback times them)
        return {"primal_launches": prof.fallback_primal_launches,
                "dual_launches": prof.fallback_dual_launches,
                "primal_bytes": prof.fallback_primal_bytes, "dual_bytes": prof.fallback_dual_bytes,
                "note": "launch durations in profiles/ (rocprofv3); not timed in this run"}
    return {"primal_launches": prof.fallback_primal_launches,
            "primal_avg_us": round(pm / prof.fallback_primal_launches * 1e3, 2),
            "primal_GBs": round(prof.fallback_primal_bytes / max(pm, 1e-12) / 1e6, 1),
            "dual_launches": prof.fallback_dual_launches,
            "dual_avg_us": round(dm / max(prof.fallback_dual_launches, 1) * 1e3, 2),
            "dual_GBs": round(prof.fallback_dual_bytes / max(dm, 1e-12) / 1e6, 1)}


def aggregate_roofline(prof, seconds):
    """All step kernels of the timed runs (FP32 engines and FP64 fallback
    engines, both half-steps): algorithmic bytes over the runs' wall time."""
    total = (prof.primal_bytes + prof.dual_bytes + prof.fallback_primal_bytes
             + prof.fallback_dual_bytes)
    gbs = total / seconds / 1e9
    return {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": total,
            "seconds": round(seconds, 3)}


# Optimal objectives of generated LPs without a certificate, from HiGHS
# (scipy.optimize.linprog, method="highs") in FP64 on the host; see
# tests/golden/highs_objectives.json.
REFERENCE_OBJECTIVES = {2: 1219334.0106920893}  # 2,000 x 2,000 transportation LP, seed 2


def measured_traffic(workload, precision, kernel, bytes_per_launch):
    rec = MEASURED_TRAFFIC.get((workload, precision, kernel))
    if rec is None or abs(rec[0] - bytes_per_launch) > 5e-3 * rec[0]:
        return None
    return round(rec[1] * bytes_per_launch / rec[0], -3)


def profile_roofline(prof, kernel_names, note, workload=None, precision=None):
    pm, dm = prof.primal_ms, prof.dual_ms
    if prof.primal_launches == 0:
        return None
    if dm >= pm:
        name, ms, nbytes, launches = kernel_names[1], dm, prof.dual_bytes, prof.dual_launches
    else:
        name, ms, nbytes, launches = kernel_names[0], pm, prof.primal_bytes, prof.primal_launches
    avg_ms = ms / launches
    bytes_per_launch = nbytes / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(workload, precision, name, bytes_per_launch), "kernel": name,
            "avg_launch_us": round(avg_ms * 1e3, 2), "bytes_per_launch": bytes_per_launch,
            "launches": launches, "note": note,
            "other_kernel": {"primal_avg_us": round(pm / max(prof.primal_launches, 1) * 1e3, 2),
                             "dual_avg_us": round(dm / max(prof.dual_launches, 1) * 1e3, 2),
                             "primal_GBs": round(prof.primal_bytes / max(pm, 1e-12) / 1e6, 1),
                             "dual_GBs": round(prof.dual_bytes / max(dm, 1e-12) / 1e6, 1)}}


def cpu_baseline_batch(problems, tol, budget_s=20.0):
    """The CPU oracle on a bounded sample: instances in batch order until the
    time budget is spent."""
    O = _oracle()
    threads = O.lib().orc_num_threads()
    p = make_params(tol, "fp64")
    O.set_exact(False)  # an ordinary FP64 CPU implementation for timing
    t0 = time.time()
    done = 0
    for prob in problems:
        O.solve(prob, p)
        done += 1
        if time.time() - t0 > budget_s:
            break
    dt = time.time() - t0
    O.set_exact(True)
    return {"value": round(done / dt, 4), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of the {len(problems)} batch LPs, one at a time, "
                      f"tol {tol:g}, {dt:.1f} s, plain FP64 sums, OpenMP"}


def _oracle():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    O.build()
    return O


def cpu_baseline_single(prob, tol, iteration_limit, gpu_iters_per_step, budget_s=20.0):
    """The CPU oracle on the same LP with the same limits, its time limit set
    to the budget: iterations/s, and the solves/s that rate gives for the GPU
    step's iteration count (the unit of `value`)."""
    O = _oracle()
    threads = O.lib().orc_num_threads()
    p = make_params(tol, "fp64", iteration_limit=iteration_limit, time_limit=budget_s)
    O.set_exact(False)
    t0 = time.time()
    out = O.solve(prob, p)
    dt = time.time() - t0
    O.set_exact(True)
    its = out.iteration_count
    ips = its / dt
    return {"value": round(ips / max(gpu_iters_per_step, 1), 6), "unit": "solves/s",
            "cores": threads, "kind": "port", "iterations_per_s": round(ips, 2),
            "sample": f"the same LP, {its} iterations ({out.termination_string}) in {dt:.1f} s "
                      f"including rescaling, plain FP64 sums, OpenMP; value = iterations/s / "
                      f"{gpu_iters_per_step:g} iterations per GPU step"}


def cpu_baseline_stream(problems, tols, budget_s=20.0):
    """The CPU oracle on the requests of the stream in arrival order, one at a
    time at each request's tolerance, until the time budget is spent."""
    O = _oracle()
    threads = O.lib().orc_num_threads()
    O.set_exact(False)
    t0 = time.time()
    done = 0
    for prob, tol in zip(problems, tols):
        O.solve(prob, make_params(tol, "fp64"))
        done += 1
        if time.time() - t0 > budget_s:
            break
    dt = time.time() - t0
    O.set_exact(True)
    return {"value": round(done / dt, 4), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of the {len(problems)} requests, one at a time at their "
                      f"tolerances, {dt:.1f} s, plain FP64 sums, OpenMP"}


def solve_resident(ctx, problems, params):
    """Loads the LPs onto the device (untimed), then solves the resident
    batch: returns (outputs, run seconds, load seconds)."""
    t0 = time.time()
    batch = ctx.batch_load(problems, params)
    t1 = time.time()
    outs = batch.run()
    t2 = time.time()
    batch.free()
    return outs, t2 - t1, t1 - t0


def run_batch(args, ctx, lib):
    import ctypes as C
    from cupdlp import _abi
    log(f"generating {args.count} LPs ...")
    t = time.time()
    problems, objectives = batch_problems(args.count)
    nnz = sum(p.constraint_matrix.nnz for p in problems)
    log(f"  {time.time() - t:.1f} s, total nnz {nnz}, total n {sum(p.num_variables for p in problems)}")
    tuning = {}
    if args.tune:
        tuning[args.precision] = tune(ctx, problems, args.precision, args.tol)
        if args.precision == "fp32" and not args.no_fp64_batch:
            tuning["fp64"] = tune(ctx, problems, "fp64", args.tol)
    params = make_params(args.tol, args.precision, time_limit=args.time_limit)
    for w in range(args.warmup):
        log(f"warmup {w}")
        ctx.batch_solve(problems, params)
    lib.pdlp_get_profile(ctx.handle, C.byref(_abi.pdlp_kernel_profile()))
    lib.pdlp_set_profiling(ctx.handle, 2 if args.profile_fallback else 1)
    lat, iters, bad, fallbacks = [], [], 0, 0
    dt, load_s = 0.0, 0.0
    for k in range(args.steps):
        outs, d_run, d_load = solve_resident(ctx, problems, params)
        dt += d_run
        load_s += d_load
        for o in outs:
            lat.append(o.solve_time_sec)
            iters.append(o.iteration_count)
            bad += int(o.termination_string != "OPTIMAL")
            fallbacks += int(o.fp64_fallback)
        log(f"step {k}: run {d_run:.2f} s, load {d_load:.2f} s")
    lib.pdlp_set_profiling(ctx.handle, 0)
    prof = _abi.pdlp_kernel_profile()
    lib.pdlp_get_profile(ctx.handle, C.byref(prof))
    value = args.steps * len(problems) / dt
    extra = {"load_inclusive_solves_per_s": round(args.steps * len(problems) / (dt + load_s), 4),
             "load_s_per_step": round(load_s / args.steps, 3)}
    if tuning:
        extra["tuning"] = tuning
    if args.dump_instances:
        with open(args.dump_instances, "w") as f:
            json.dump([{"i": i, "n": p.num_variables, "nnz": int(p.constraint_matrix.nnz),
                        "iterations": o.iteration_count, "attempts": o.total_attempts,
                        "rounds": o.refinement_rounds, "fallback": int(o.fp64_fallback),
                        "latency_s": round(o.solve_time_sec, 4)}
                       for i, (p, o) in enumerate(zip(problems, outs))], f)
    # error of every returned point of the last timed step
    extra.update(accuracy(problems, outs, objectives))
    extra["aggregate_step_kernels"] = aggregate_roofline(prof, dt)
    extra["fallback_kernels"] = fallback_kernels(prof)
    if args.precision == "fp32" and not args.no_fp64_batch:
        # the same batch in plain FP64, batched (one timed solve)
        o64, d64, _ = solve_resident(ctx, problems,
                                     make_params(args.tol, "fp64", time_limit=args.time_limit))
        extra["fp64_batched_solves_per_s"] = round(len(problems) / d64, 4)
        extra["fp64_batched_p99_latency_s"] = round(float(np.percentile([o.solve_time_sec for o in o64], 99)), 4)
        if args.dump_instances:
            with open(args.dump_instances + ".fp64.json", "w") as f:
                json.dump([{"i": i, "n": p.num_variables, "nnz": int(p.constraint_matrix.nnz),
                            "iterations": o.iteration_count, "attempts": o.total_attempts,
                            "latency_s": round(o.solve_time_sec, 4)}
                           for i, (p, o) in enumerate(zip(problems, o64))], f)
        a64 = accuracy(problems, o64, objectives)
        extra["fp64_batched_max_rel_kkt"] = a64["max_rel_kkt"]
        extra["fp64_batched_max_rel_obj_error_vs_certificate"] = a64["max_rel_obj_error_vs_certificate"]
        log(f"fp64 batched: {d64:.2f} s")
    # plain FP64 one at a time (each LP loaded, then its solve timed, as
    # above), the whole batch unless the budget runs out first; its latency
    # is the completion time with the batch submitted at once and solved in
    # order, its per-solve time beside it
    p64 = make_params(args.tol, "fp64")
    t1 = time.time()
    ser_times = []
    for prob in problems:
        _, d_run, _ = solve_resident(ctx, [prob], p64)
        ser_times.append(d_run)
        if time.time() - t1 > args.serial_budget:
            break
    nser, ser_s = len(ser_times), float(np.sum(ser_times))
    ser = nser / ser_s
    extra.update({
        "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
        "not_optimal": bad,
        "fp64_fallbacks": fallbacks,
        "median_iterations": int(np.median(iters)),
        "fp64_serial_solves_per_s": round(ser, 3),
        "fp64_serial_sample": nser,
        "fp64_serial_p99_latency_s": round(float(np.percentile(np.cumsum(ser_times), 99)), 4),
        "fp64_serial_p50_latency_s": round(float(np.percentile(np.cumsum(ser_times), 50)), 4),
        "fp64_serial_p99_solve_s": round(float(np.percentile(ser_times, 99)), 4),
        "speedup_vs_fp64_serial": round(value / ser, 3),
    })
    if args.serve_count > 0:
        extra["served"] = served_leg(args, ctx)
    roof = profile_roofline(prof, ("k_primal", "k_dual"),
                            f"{args.count}-LP batch, {args.precision}, tol {args.tol:g}",
                            "batch", args.precision)
    form = "FP32 + FP64 refinement" if args.precision == "fp32" else "plain FP64"
    cfg = {"workload": f"batch of {args.count} random LPs (seeds 0..{args.count - 1}, n log-uniform "
                       f"500..200000, m=n/2, 8 nnz/col), {form}, batched",
           "tolerance": args.tol, "precision": args.precision, "total_nnz": nnz}
    return value, dt, cfg, extra, roof, problems


def serve_stream(count, nmax, seed=4):
    """A request stream after the served configuration: 90 % random LPs with
    n log-uniform in [1,000, nmax] (m = n/2, 8 nonzeros per column), 10 %
    transportation LPs with 200..1,000 sources and sinks; tolerance 1e-4
    (70 %) or 1e-8 (30 %).  Per-request seeds come from the stream seed."""
    from cupdlp.generators import random_lp, transportation_lp
    rng = np.random.default_rng(seed)
    probs, tols = [], []
    for k in range(count):
        if k and k % 32 == 0:
            log(f"  {k} of {count} requests generated")
        s = int(rng.integers(0, 2**31 - 1))
        if rng.random() < 0.9:
            n = int(np.exp(rng.uniform(np.log(1000), np.log(nmax))))
            probs.append(random_lp(n=n, m=n // 2, seed=s)[0])
        else:
            a, b = (int(v) for v in rng.integers(200, 1001, size=2))
            probs.append(transportation_lp(a, b, seed=s))
        tols.append(1e-4 if rng.random() < 0.7 else 1e-8)
    return probs, tols


def serve_once(ctx, probs, tols, rate, precision, hbm_budget_gb, request_time_limit=0.0):
    """Streams the requests through the request loop with Poisson arrivals
    at `rate`, each with its own time limit when request_time_limit > 0
    (seconds from submission): returns (seconds, latencies, results not
    OPTIMAL, results that ended at their time limit)."""
    from cupdlp.solver import Server
    rng = np.random.default_rng(5)
    gaps = rng.exponential(1.0 / rate, size=len(probs)) if rate > 0 else np.zeros(len(probs))
    arrive = np.cumsum(gaps) - gaps[0]
    srv = Server(ctx, make_params(1e-4, precision), hbm_budget_bytes=hbm_budget_gb * 1e9)
    lat, bad, timed_out, done = [], 0, 0, 0
    limit = request_time_limit if request_time_limit > 0 else None
    nxt = 0
    t0 = time.time()
    last_log = t0
    while done < len(probs):
        if time.time() - last_log > 30.0:
            last_log = time.time()
            log(f"  {last_log - t0:.0f} s: {nxt} submitted, {done} done")
        now = time.time() - t0
        while nxt < len(probs) and arrive[nxt] <= now:
            srv.submit(probs[nxt], tols[nxt], time_sec_limit=limit)
            nxt += 1
        if done < nxt:
            for f in srv.poll():
                o = srv.result(f)
                lat.append(o.solve_time_sec)
                bad += int(o.termination_string != "OPTIMAL")
                timed_out += int(o.termination_string == "TIME_LIMIT")
                done += 1
        elif nxt < len(probs):
            time.sleep(max(0.0, min(arrive[nxt] - now, 0.05)))
    dt = time.time() - t0
    srv.close()
    return dt, lat, bad, timed_out


def serial_stream(ctx, probs, tols, budget, request_time_limit=0.0):
    """Plain FP64 one request at a time, in arrival order, until the budget
    is spent, each within request_time_limit seconds when that is > 0:
    (solves/s, requests solved)."""
    t1 = time.time()
    nser = 0
    limit = request_time_limit if request_time_limit > 0 else float("inf")
    for p, t in zip(probs, tols):
        ctx.solve(p, make_params(t, "fp64", time_limit=limit))
        nser += 1
        if time.time() - t1 > budget:
            break
    return nser / (time.time() - t1), nser


def served_leg(args, ctx):
    """A bounded served stream beside the batch line: args.serve_count
    requests of the served mix (random LPs n <= args.serve_nmax, 10 %
    transportation), Poisson arrivals at args.rate, in the batch line's
    precision, each request with a time limit of args.serve_time_limit
    seconds from its submission (0: none), against plain FP64 one at a time
    on the same requests with the same limit.  The default stream holds one
    LP that needs 470,592 iterations to 1e-8 (21 s alone): without a limit
    it alone sets the stream's makespan."""
    log(f"served leg: generating {args.serve_count} requests ...")
    probs, tols = serve_stream(args.serve_count, args.serve_nmax)
    lim = args.serve_time_limit
    dt, lat, bad, timed_out = serve_once(ctx, probs, tols, args.rate, args.precision,
                                         args.hbm_budget, lim)
    ser, nser = serial_stream(ctx, probs, tols, args.serial_budget / 2, lim)
    value = len(probs) / dt
    log(f"served leg: {value:.2f} solves/s, FP64 one at a time {ser:.2f}")
    return {"requests": len(probs), "completed": len(lat), "not_optimal": bad,
            "time_limited": timed_out, "request_time_limit_s": lim or None,
            "solves_per_s": round(value, 3), "seconds": round(dt, 2),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "fp64_serial_solves_per_s": round(ser, 3), "fp64_serial_sample": nser,
            "speedup_vs_fp64_serial": round(value / ser, 3),
            "workload": f"{len(probs)} requests: 90% random LPs n log-uniform 1000..{args.serve_nmax}, "
                        f"10% transportation 200..1000 per side; tolerance 1e-4 (70%) / 1e-8 "
                        f"(30%); Poisson arrivals at {args.rate:g}/s; {args.precision}; HBM budget "
                        f"{args.hbm_budget:g} GB"
                        + (f"; per-request time limit {lim:g} s" if lim else "")}


def run_serve(args, ctx, lib):
    log(f"generating a stream of {args.count} requests ...")
    probs, tols = serve_stream(args.count, args.serve_nmax)
    for w in range(args.warmup):  # warm-up: the first few requests alone
        for p, t in list(zip(probs, tols))[:4]:
            ctx.solve(p, make_params(t, args.precision))
    dt, lat, bad, timed_out = serve_once(ctx, probs, tols, args.rate, args.precision,
                                         args.hbm_budget, args.request_time_limit)
    value = len(probs) / dt
    ser, nser = serial_stream(ctx, probs, tols, args.serial_budget, args.request_time_limit)
    extra = {"p50_latency_s": round(float(np.percentile(lat, 50)), 4),
             "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
             "not_optimal": bad, "time_limited": timed_out,
             "request_time_limit_s": args.request_time_limit or None,
             "fp64_serial_solves_per_s": round(ser, 3),
             "fp64_serial_sample": nser, "speedup_vs_fp64_serial": round(value / ser, 3)}
    cfg = {"workload": f"served stream of {len(probs)} requests (90% random LPs n log-uniform "
                       f"1000..{args.serve_nmax}, 10% transportation 200..1000 per side; tolerance "
                       f"1e-4 (70%) / 1e-8 (30%)), Poisson arrivals at {args.rate:g}/s",
           "precision": args.precision, "hbm_budget_GB": args.hbm_budget}
    extra["_tols"] = tols
    return value, dt, cfg, extra, None, probs


SINGLE_LP_NAMES = {0: "n=2,000 random LP", 1: "n=1M Zipf-degree LP",
                   2: "2,000 x 2,000 transportation LP"}


def run_single(args, ctx, lib, which):
    import ctypes as C
    from cupdlp import _abi
    from cupdlp.generators import config_lp
    name = SINGLE_LP_NAMES[which]
    log(f"generating the {name} ...")
    prob, cert = config_lp(which)
    tuning = tune(ctx, [prob], args.precision, args.tol) if args.tune else None
    params = make_params(args.tol, args.precision, iteration_limit=args.iters,
                         time_limit=args.time_limit)
    for w in range(args.warmup):
        ctx.solve(prob, params)
    lib.pdlp_set_profiling(ctx.handle, 1)
    its, dt = 0, 0.0
    for k in range(args.steps):
        (out,), d_run, _ = solve_resident(ctx, [prob], params)
        its += out.iteration_count
        dt += d_run
    lib.pdlp_set_profiling(ctx.handle, 0)
    prof = _abi.pdlp_kernel_profile()
    lib.pdlp_get_profile(ctx.handle, C.byref(prof))
    roof = profile_roofline(prof, ("k_primal", "k_dual"), f"{name}, {args.precision}",
                            f"c{which}", args.precision)
    # the same solve once more without the kernel events of the timed steps,
    # which cost a launch-bound (small) LP a large share of its iteration
    (out_u,), d_u, _ = solve_resident(ctx, [prob], params)
    n, m, nnz = prob.num_variables, prob.num_constraints, prob.constraint_matrix.nnz
    cfg = {"workload": f"{name}: n={n} m={m} nnz={nnz}", "iteration_limit": args.iters,
           "tolerance": args.tol, "precision": args.precision}
    ci = out.iteration_stats[-1].convergence_information[0]
    extra = {"iterations_per_s": round(its / dt, 1), "iterations": its,
             "iterations_per_s_unprofiled": round(out_u.iteration_count / d_u, 1),
             "termination": out.termination_string, "solve_time_s": round(out.solve_time_sec, 3),
             "relative_l2_primal_residual": ci.relative_l2_primal_residual,
             "relative_l2_dual_residual": ci.relative_l2_dual_residual,
             "relative_optimality_gap": ci.relative_optimality_gap,
             "primal_objective": ci.primal_objective,
             "refinement_rounds": out.refinement_rounds, "fp64_fallback": out.fp64_fallback}
    if tuning:
        extra["tuning"] = tuning
    ref = cert.objective if cert is not None else REFERENCE_OBJECTIVES.get(which)
    if ref is not None:
        extra["objective_reference"] = ref
        extra["objective_rel_error"] = abs(ci.primal_objective - ref) / max(1.0, abs(ref))
    return args.steps / dt, dt, cfg, extra, roof, [prob]


def run_copy(args):
    """A plain device-to-device copy of a known size (1 GiB, HBM-resident),
    the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    nbytes = 1 << 30
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda:0").uniform_()
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    print(json.dumps({"workload": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "avg_ms": round(ms, 4), "GBs": round(gbs, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("batch", "serve", "c1", "c2", "c0", "copy"), default="batch")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-nmax", type=int, default=200000, help="serve: largest random-LP n")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serve: HBM budget (GB)")
    ap.add_argument("--request-time-limit", type=float, default=0.0,
                    help="serve: each request's own time limit in seconds from its "
                         "submission (pdlp_serve_submit_limits; 0 = none)")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default=None,
                    help="default: fp32 (FP32 + FP64 refinement) for the batch, fp64 otherwise")
    ap.add_argument("--no-fp64-batch", action="store_true",
                    help="batch: skip the plain FP64 batched solve reported beside the FP32 one")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--batch-groups", type=int, default=0,
                    help="batch: size groups run side by side (0 = the library's default)")
    ap.add_argument("--refine-inner-tol", type=float, default=0.0,
                    help="FP32 + refinement: tolerance of the first FP32 solve and the tightest "
                         "of the correction solves (0 = the library's default, 1e-4)")
    ap.add_argument("--refine-early-fallback", type=int, default=0,
                    help="FP32 + refinement, batches: iterations into a correction solve after "
                         "which the FP64 fallback starts beside it (0 = never, the default)")
    ap.add_argument("--refine-cap", type=int, default=0,
                    help="FP32 + refinement: iteration cap of each correction solve "
                         "(0 = the library default)")
    ap.add_argument("--slice-occupancy", default="",
                    help="P[,D]: most waves per SIMD of the primal [and dual] slice kernels "
                         "(0: the library's defaults; 8: no cap; empty: the tuner's choice, "
                         "or the defaults without tuning)")
    ap.add_argument("--refine-continue", type=int, default=-1,
                    help="FP32 + refinement: a capped correction solve that cut the error this "
                         "many times continues the refinement (0: stall at once; -1: the "
                         "library default)")
    ap.add_argument("--refine-warm-start", type=int, default=-1,
                    help="FP32 + refinement: correction solves start from the previous "
                         "solve's primal weight (1), step size (2) or both (3)")
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--serial-budget", type=float, default=60.0,
                    help="seconds of plain FP64 one-at-a-time solves (the batch: stops early "
                         "only if its LPs take longer; the served leg: half of it)")
    ap.add_argument("--tune", type=int, default=1,
                    help="batch and single-LP workloads: pick the step kernels' tile "
                         "parameters with pdlp_tune before anything is timed (0 = defaults)")
    ap.add_argument("--profile-fallback", action="store_true",
                    help="batch: time the FP64 fallback engines' launches too (events on "
                         "every launch; slows the run)")
    ap.add_argument("--dump-instances", default="",
                    help="batch: write each instance's size, iterations and latency (last "
                         "timed step) to this JSON file")
    ap.add_argument("--serve-time-limit", type=float, default=10.0,
                    help="served leg of the batch line: time limit of each request (s, from "
                         "its submission; 0: none)")
    ap.add_argument("--serve-count", type=int, default=160,
                    help="batch: requests of the served leg reported beside it (0 = none)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--time-limit", type=float, default=float("inf"),
                    help="per-solve time limit (a guard; instances that hit it count as not optimal)")
    args = ap.parse_args()
    global BATCH_GROUPS, REFINE_INNER_TOL, REFINE_EARLY, REFINE_WARM
    REFINE_WARM = args.refine_warm_start
    global REFINE_CAP, REFINE_CONTINUE, OCCUPANCY
    REFINE_CAP = args.refine_cap
    REFINE_CONTINUE = args.refine_continue
    if args.slice_occupancy:
        occ = [int(v) for v in args.slice_occupancy.split(",")]
        OCCUPANCY = (occ[0], occ[-1])
    BATCH_GROUPS = args.batch_groups
    REFINE_INNER_TOL = args.refine_inner_tol
    REFINE_EARLY = args.refine_early_fallback
    if args.precision is None:
        args.precision = "fp32" if args.workload == "batch" else "fp64"
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs on a single device)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        run_copy(args)
        return
    from cupdlp.solver import Context
    ctx = Context(0)
    lib = ctx.lib
    if args.workload == "batch":
        value, dt, cfg, extra, roof, problems = run_batch(args, ctx, lib)
        metric_unit = "solves/s"
    elif args.workload == "serve":
        value, dt, cfg, extra, roof, problems = run_serve(args, ctx, lib)
        metric_unit = "solves/s"
    else:
        value, dt, cfg, extra, roof, problems = run_single(args, ctx, lib, int(args.workload[1]))
        metric_unit = "solves/s"
    cpu = None
    tols = extra.pop("_tols", None)
    if not args.no_cpu_baseline:
        if args.workload == "batch":
            cpu = cpu_baseline_batch(problems, args.tol, args.cpu_budget)
        elif args.workload == "serve":
            cpu = cpu_baseline_stream(problems, tols, args.cpu_budget)
        else:
            cpu = cpu_baseline_single(problems[0], args.tol, args.iters,
                                      extra["iterations"] / args.steps, args.cpu_budget)
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": metric_unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded generators with certified optima)", "config": cfg,
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
