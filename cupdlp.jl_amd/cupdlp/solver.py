"""``optimize(params, problem)`` over the C ABI of libpdlp_hip.so.

Mirrors cuPDLP.optimize (src/primal_dual_hybrid_gradient_gpu.jl:457-798):
same arguments, same SaddlePointOutput, same error behaviour (invalid input
raises, like the reference's ``error()`` from ``validate``).
"""
import ctypes as C
import threading

import numpy as np

from . import _abi
from .types import (AdaptiveStepsizeParams, ConvergenceInformation, InfeasibilityInformation,
                    IterationStats, PdhgParameters, PointType, QuadraticProgrammingProblem,
                    RestartChoice, SaddlePointOutput, TerminationReason,
                    termination_reason_to_string)


class PdlpError(RuntimeError):
    """Raised for invalid input (code > 0) or a device failure (code < 0)."""

    def __init__(self, code, msg):
        super().__init__(f"pdlp error {code}: {msg}")
        self.code = code


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def to_c_lp(problem: QuadraticProgrammingProblem):
    """Builds a borrowed pdlp_lp view (1-based CSC, Int64) and the arrays it
    points to (kept alive by the returned tuple)."""
    A = problem.constraint_matrix
    m, n = A.shape
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int64) + 1
    rowval = np.ascontiguousarray(A.indices, dtype=np.int64) + 1
    nzval = np.ascontiguousarray(A.data, dtype=np.float64)
    c = np.ascontiguousarray(problem.objective_vector, dtype=np.float64)
    lb = np.ascontiguousarray(problem.variable_lower_bound, dtype=np.float64)
    ub = np.ascontiguousarray(problem.variable_upper_bound, dtype=np.float64)
    b = np.ascontiguousarray(problem.right_hand_side, dtype=np.float64)
    if len(c) != n or len(lb) != n or len(ub) != n or len(b) != m:
        raise PdlpError(1, "vector lengths do not match the constraint matrix "
                           "(src/preprocess.jl:8-74)")
    keep = (colptr, rowval, nzval, c, lb, ub, b)
    lp = _abi.pdlp_lp(n=n, m=m, num_eq=problem.num_equalities, nnz=int(A.nnz),
                      colptr=_ptr(colptr, C.c_int64), rowval=_ptr(rowval, C.c_int64),
                      nzval=_ptr(nzval, C.c_double), c=_ptr(c, C.c_double),
                      lb=_ptr(lb, C.c_double), ub=_ptr(ub, C.c_double),
                      b=_ptr(b, C.c_double), c0=problem.objective_constant)
    return lp, keep


def to_c_params(p: PdhgParameters) -> _abi.pdlp_params:
    cp = _abi.pdlp_params()
    cp.l_inf_ruiz_iterations = int(p.l_inf_ruiz_iterations)
    cp.l2_norm_rescaling = int(bool(p.l2_norm_rescaling))
    cp.has_pock_chambolle_alpha = 0 if p.pock_chambolle_alpha is None else 1
    cp.pock_chambolle_alpha = 0.0 if p.pock_chambolle_alpha is None else float(p.pock_chambolle_alpha)
    cp.primal_importance = float(p.primal_importance)
    cp.scale_invariant_initial_primal_weight = int(bool(p.scale_invariant_initial_primal_weight))
    cp.record_iteration_stats = int(bool(p.record_iteration_stats))
    cp.verbosity = int(p.verbosity)
    cp.termination_evaluation_frequency = int(p.termination_evaluation_frequency)
    if isinstance(p.step_size_policy_params, AdaptiveStepsizeParams):
        cp.step_size_policy = 0
        cp.reduction_exponent = float(p.step_size_policy_params.reduction_exponent)
        cp.growth_exponent = float(p.step_size_policy_params.growth_exponent)
    else:
        cp.step_size_policy = 1
    tc = p.termination_criteria
    cp.termination_criteria.optimality_norm = int(tc.optimality_norm)
    cp.termination_criteria.iteration_limit = int(min(tc.iteration_limit, 2**31 - 1))
    cp.termination_criteria.eps_optimal_absolute = float(tc.eps_optimal_absolute)
    cp.termination_criteria.eps_optimal_relative = float(tc.eps_optimal_relative)
    cp.termination_criteria.eps_primal_infeasible = float(tc.eps_primal_infeasible)
    cp.termination_criteria.eps_dual_infeasible = float(tc.eps_dual_infeasible)
    cp.termination_criteria.time_sec_limit = float(tc.time_sec_limit)
    cp.termination_criteria.kkt_matrix_pass_limit = float(tc.kkt_matrix_pass_limit)
    rp = p.restart_params
    cp.restart_params.restart_scheme = int(rp.restart_scheme)
    cp.restart_params.restart_to_current_metric = int(rp.restart_to_current_metric)
    cp.restart_params.restart_frequency_if_fixed = int(rp.restart_frequency_if_fixed)
    cp.restart_params.artificial_restart_threshold = float(rp.artificial_restart_threshold)
    cp.restart_params.sufficient_reduction_for_restart = float(rp.sufficient_reduction_for_restart)
    cp.restart_params.necessary_reduction_for_restart = float(rp.necessary_reduction_for_restart)
    cp.restart_params.primal_weight_update_smoothing = float(rp.primal_weight_update_smoothing)
    cp.precision = int(p.precision)
    cp.refine_max_outer = int(p.refine_max_outer)
    cp.refine_inner_tolerance = float(p.refine_inner_tolerance)
    cp.refine_max_scale = float(p.refine_max_scale)
    cp.device_preconditioning = int(bool(p.device_preconditioning))
    cp.refine_inner_iteration_limit = int(p.refine_inner_iteration_limit)
    cp.batch_groups = int(p.batch_groups)
    cp.refine_early_fallback = int(p.refine_early_fallback)
    cp.debug_numerical_error_attempt = int(p.debug_numerical_error_attempt)
    cp.refine_warm_start = int(p.refine_warm_start)
    cp.lane_budget = int(p.lane_budget)
    cp.slice_unroll = int(p.slice_unroll)
    cp.slice_waves = int(p.slice_waves)
    cp.refine_continue_progress = int(p.refine_continue_progress)
    cp.primal_occupancy = int(p.primal_occupancy)
    cp.dual_occupancy = int(p.dual_occupancy)
    return cp


_CI_FIELDS = [f for f, _ in _abi.pdlp_convergence_information._fields_ if f not in ("candidate_type", "pad_")]
_II_FIELDS = [f for f, _ in _abi.pdlp_infeasibility_information._fields_ if f not in ("candidate_type", "pad_")]


def from_c_stats(s: _abi.pdlp_iteration_stats) -> IterationStats:
    ci = ConvergenceInformation(PointType(s.convergence_information.candidate_type),
                                *[getattr(s.convergence_information, f) for f in _CI_FIELDS])
    ii = InfeasibilityInformation(PointType(s.infeasibility_information.candidate_type),
                                  *[getattr(s.infeasibility_information, f) for f in _II_FIELDS])
    return IterationStats(
        iteration_number=s.iteration_number, convergence_information=[ci],
        infeasibility_information=[ii],
        cumulative_kkt_matrix_passes=s.cumulative_kkt_matrix_passes,
        cumulative_rejected_steps=s.cumulative_rejected_steps,
        cumulative_time_sec=s.cumulative_time_sec, restart_used=RestartChoice(s.restart_used),
        step_size=s.step_size, primal_weight=s.primal_weight,
        method_specific_stats={
            "time_spent_doing_basic_algorithm": s.time_spent_doing_basic_algorithm,
            "total_attempts": float(s.total_attempts)})


def run_solve(call, problem: QuadraticProgrammingProblem, params: PdhgParameters,
              want_current=False):
    """Shared marshalling: ``call(lp, params, result, cb, user) -> int``."""
    lp, keep = to_c_lp(problem)
    cp = to_c_params(params)
    n, m = problem.num_variables, problem.num_constraints
    x = np.zeros(max(n, 1)); y = np.zeros(max(m, 1))
    xc = np.zeros(max(n, 1)) if want_current else None
    yc = np.zeros(max(m, 1)) if want_current else None
    res = _abi.pdlp_result()
    res.x = _ptr(x, C.c_double); res.y = _ptr(y, C.c_double)
    if want_current:
        res.x_current = _ptr(xc, C.c_double); res.y_current = _ptr(yc, C.c_double)
    stats = []

    def _cb(sp, _user):
        stats.append(from_c_stats(sp.contents))

    cb = _abi.STATS_CB(_cb)
    rc = call(C.byref(lp), C.byref(cp), C.byref(res), cb, None)
    del keep
    if rc != 0:
        return rc, None
    final = from_c_stats(res.final_stats)
    if not stats or stats[-1].iteration_number != final.iteration_number:
        stats.append(final)
    else:
        stats[-1] = final
    reason = TerminationReason(res.termination_reason)
    out = SaddlePointOutput(
        primal_solution=x[:n].copy(), dual_solution=y[:m].copy(), termination_reason=reason,
        termination_string=termination_reason_to_string(reason),
        iteration_count=int(res.iteration_count), iteration_stats=stats,
        solve_time_sec=res.solve_time_sec, preprocess_time_sec=res.preprocess_time_sec,
        total_attempts=int(res.total_attempts),
        refinement_rounds=int(res.refinement_rounds), fp64_fallback=bool(res.fp64_fallback),
        primal_current=None if xc is None else xc[:n].copy(),
        dual_current=None if yc is None else yc[:m].copy())
    return 0, out


class Context:
    """One HIP stream + device pool on one GPU (pdlp_ctx)."""

    def __init__(self, device=0):
        self.lib = _abi.load()
        h = C.c_void_p()
        rc = self.lib.pdlp_ctx_create(int(device), C.byref(h))
        if rc != 0:
            raise PdlpError(rc, self.lib.pdlp_last_error(None).decode())
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.pdlp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.lib.pdlp_last_error(self.handle).decode()

    def solve(self, problem, params, want_current=False):
        def call(lp, cp, res, cb, user):
            return self.lib.pdlp_solve(self.handle, lp, cp, res, cb, user)
        rc, out = run_solve(call, problem, params, want_current)
        if rc != 0:
            raise PdlpError(rc, self.error())
        return out

    def solve_noalloc(self, problem, params, want_current=False):
        """``solve`` through pdlp_solve_noalloc: no device allocation (a
        fixed 128 MiB region per device), FP64 only; calls on one device
        must not overlap."""
        def call(lp, cp, res, cb, user):
            return self.lib.pdlp_solve_noalloc(self.handle, lp, cp, res, cb, user)
        rc, out = run_solve(call, problem, params, want_current)
        if rc != 0:
            raise PdlpError(rc, self.error())
        return out

    def tune(self, problems, params, trial_iterations=100):
        """pdlp_tune: the step-kernel tile parameters for these LPs (batched
        with ``params``) from timed trials; a dict whose lane_budget,
        slice_unroll, slice_waves, primal_occupancy and dual_occupancy go into
        PdhgParameters.  Cached per LP shape on this context."""
        k = len(problems)
        lps = (_abi.pdlp_lp * k)()
        keeps = []
        for i, p in enumerate(problems):
            lp, keep = to_c_lp(p)
            lps[i] = lp
            keeps.append(keep)
        cp = to_c_params(params)
        out = _abi.pdlp_tuning()
        rc = self.lib.pdlp_tune(self.handle, k, lps, C.byref(cp), int(trial_iterations),
                                C.byref(out))
        del keeps
        if rc != 0:
            raise PdlpError(rc, self.error())
        return {f: getattr(out, f) for f, _ in _abi.pdlp_tuning._fields_}

    def batch_solve(self, problems, params):
        k = len(problems)
        lps = (_abi.pdlp_lp * k)()
        keeps, outs, results = [], [], (_abi.pdlp_result * k)()
        for i, p in enumerate(problems):
            lp, keep = to_c_lp(p)
            lps[i] = lp
            n, m = p.num_variables, p.num_constraints
            x = np.zeros(max(n, 1)); y = np.zeros(max(m, 1))
            keeps.append((keep, x, y))
            results[i].x = _ptr(x, C.c_double); results[i].y = _ptr(y, C.c_double)
        cp = to_c_params(params)
        rc = self.lib.pdlp_batch_solve(self.handle, k, lps, C.byref(cp), results)
        if rc != 0:
            raise PdlpError(rc, self.error())
        for i, p in enumerate(problems):
            (_, x, y) = keeps[i]
            outs.append(_output_from_result(results[i], x, y, p.num_variables, p.num_constraints))
        return outs

    def batch_load(self, problems, params):
        """Copies the LPs to the device without solving them (pdlp_batch_load);
        ``run()`` on the result solves them as ``batch_solve`` would."""
        return LoadedBatch(self, problems, params)


class LoadedBatch:
    """LPs resident on the device (pdlp_batch_load), solved once by ``run``
    (pdlp_batch_run); ``solve_time_sec`` of its results counts from the start
    of the run.  No counterpart in the reference."""

    def __init__(self, ctx, problems, params):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = None
        k = len(problems)
        lps = (_abi.pdlp_lp * k)()
        keeps = []
        for i, p in enumerate(problems):
            lp, keep = to_c_lp(p)
            lps[i] = lp
            keeps.append(keep)
        cp = to_c_params(params)
        h = C.c_void_p()
        rc = self.lib.pdlp_batch_load(ctx.handle, k, lps, C.byref(cp), C.byref(h))
        del keeps  # the library keeps no pointer into the caller's arrays
        if rc != 0:
            raise PdlpError(rc, ctx.error())
        self.handle = h
        self.sizes = [(p.num_variables, p.num_constraints) for p in problems]

    def run(self):
        k = len(self.sizes)
        results = (_abi.pdlp_result * k)()
        bufs = []
        for i, (n, m) in enumerate(self.sizes):
            x = np.zeros(max(n, 1)); y = np.zeros(max(m, 1))
            bufs.append((x, y))
            results[i].x = _ptr(x, C.c_double); results[i].y = _ptr(y, C.c_double)
        rc = self.lib.pdlp_batch_run(self.handle, results)
        if rc != 0:
            raise PdlpError(rc, self.ctx.error())
        return [_output_from_result(results[i], x, y, n, m)
                for i, ((x, y), (n, m)) in enumerate(zip(bufs, self.sizes))]

    def free(self):
        if self.handle:
            self.lib.pdlp_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _output_from_result(r, x, y, n, m):
    reason = TerminationReason(r.termination_reason)
    return SaddlePointOutput(
        primal_solution=x[:n].copy(), dual_solution=y[:m].copy(),
        termination_reason=reason, termination_string=termination_reason_to_string(reason),
        iteration_count=int(r.iteration_count), iteration_stats=[from_c_stats(r.final_stats)],
        solve_time_sec=r.solve_time_sec, preprocess_time_sec=r.preprocess_time_sec,
        total_attempts=int(r.total_attempts),
        refinement_rounds=int(r.refinement_rounds), fp64_fallback=bool(r.fp64_fallback))


class Server:
    """Continuous-batching request loop (pdlp_serve_*; no counterpart in the
    reference).  Requests are admitted by HBM footprint and slot count as
    they arrive, stepped together with the running ones and retired when
    they finish; each follows ``optimize`` with eps_optimal_absolute =
    eps_optimal_relative = its tolerance, so its result equals a lone solve
    with those criteria.  ``solve_time_sec`` of a result is its latency from
    ``submit``."""

    def __init__(self, ctx: Context, params: PdhgParameters, hbm_budget_bytes=0.0, max_slots=0):
        self.ctx = ctx
        self.lib = ctx.lib
        self._params = to_c_params(params)
        h = C.c_void_p()
        rc = self.lib.pdlp_serve_open(ctx.handle, C.byref(self._params), float(hbm_budget_bytes),
                                      int(max_slots), C.byref(h))
        if rc != 0:
            raise PdlpError(rc, ctx.error())
        self.handle = h
        self._shape = {}

    def submit(self, problem: QuadraticProgrammingProblem, tolerance: float,
               time_sec_limit=None, iteration_limit=None) -> int:
        """Queues a request; optional limits of its own apply on top of the
        server's (time counted from submission, queueing included)."""
        lp, keep = to_c_lp(problem)
        rid = C.c_int64()
        if time_sec_limit is None and iteration_limit is None:
            rc = self.lib.pdlp_serve_submit(self.handle, C.byref(lp), float(tolerance),
                                            C.byref(rid))
        else:
            rc = self.lib.pdlp_serve_submit_limits(
                self.handle, C.byref(lp), float(tolerance),
                float("inf") if time_sec_limit is None else float(time_sec_limit),
                -1 if iteration_limit is None else int(min(iteration_limit, 2**31 - 1)),
                C.byref(rid))
        del keep
        if rc != 0:
            raise PdlpError(rc, self.ctx.error())
        self._shape[rid.value] = (problem.num_variables, problem.num_constraints)
        return rid.value

    def poll(self, max_ids=4096):
        """Advances the device work by one round; returns the ids finished."""
        ids = (C.c_int64 * max_ids)()
        k = C.c_int32()
        rc = self.lib.pdlp_serve_poll(self.handle, ids, int(max_ids), C.byref(k))
        if rc != 0:
            raise PdlpError(rc, self.ctx.error())
        return [ids[i] for i in range(k.value)]

    def cancel(self, rid: int):
        """Cancels a queued or running request: it finishes with reason
        OTHER and no solution, reported by the next ``poll``."""
        rc = self.lib.pdlp_serve_cancel(self.handle, int(rid))
        if rc != 0:
            raise PdlpError(rc, self.ctx.error())

    def result(self, rid: int) -> SaddlePointOutput:
        n, m = self._shape.pop(rid)
        x = np.zeros(max(n, 1)); y = np.zeros(max(m, 1))
        r = _abi.pdlp_result()
        r.x = _ptr(x, C.c_double); r.y = _ptr(y, C.c_double)
        rc = self.lib.pdlp_serve_result(self.handle, int(rid), C.byref(r))
        if rc != 0:
            raise PdlpError(rc, self.ctx.error())
        return _output_from_result(r, x, y, n, m)

    def close(self):
        if self.handle:
            self.lib.pdlp_serve_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_lock = threading.Lock()
_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def optimize(params: PdhgParameters, original_problem: QuadraticProgrammingProblem,
             want_current=False) -> SaddlePointOutput:
    """cuPDLP.optimize on the GPU (src/primal_dual_hybrid_gradient_gpu.jl:457-798)."""
    return default_context().solve(original_problem, params, want_current)


def default_params() -> PdhgParameters:
    """Defaults of cuPDLP.Optimizer() (src/MOI_wrapper.jl:58-78)."""
    return PdhgParameters()
