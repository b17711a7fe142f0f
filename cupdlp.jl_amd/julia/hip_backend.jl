# hip_backend.jl — the HIP (MI355X) backend of the cuPDLP module.
#
# Include this file in `module cuPDLP` (src/cuPDLP.jl) in place of the four
# CUDA files cpu_to_gpu.jl, iteration_stats_utils_gpu.jl, saddle_point_gpu.jl
# and primal_dual_hybrid_gradient_gpu.jl, after termination.jl and
# solve_log.jl, and drop `using CUDA`.  It defines the same public names those
# files define for the rest of the module and for MOI_wrapper.jl
# (AdaptiveStepsizeParams, ConstantStepsizeParams, PdhgParameters,
# RestartScheme, RestartToCurrentMetric, RestartParameters,
# construct_restart_parameters, SaddlePointOutput, optimize), with the same
# fields and argument meaning, and runs the solve through the C ABI of
# libpdlp_hip.so (include/pdlp_hip.h).
#
# The library path comes from ENV["PDLP_HIP_LIB"], else the package's
# lib/libpdlp_hip.so next to this directory.

const LIBPDLP = get(ENV, "PDLP_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libpdlp_hip.so"))

# ----------------------------------------------------------------- public types
# (src/primal_dual_hybrid_gradient_gpu.jl:1-24, src/saddle_point_gpu.jl:3-34,
# 336-403 — same names, fields and field order)

struct AdaptiveStepsizeParams
    reduction_exponent::Float64
    growth_exponent::Float64
end

struct ConstantStepsizeParams end

@enum RestartScheme NO_RESTARTS FIXED_FREQUENCY ADAPTIVE_KKT
@enum RestartToCurrentMetric NO_RESTART_TO_CURRENT KKT_GREEDY

mutable struct RestartParameters
    restart_scheme::RestartScheme
    restart_to_current_metric::RestartToCurrentMetric
    restart_frequency_if_fixed::Int64
    artificial_restart_threshold::Float64
    sufficient_reduction_for_restart::Float64
    necessary_reduction_for_restart::Float64
    primal_weight_update_smoothing::Float64
end

function construct_restart_parameters(
    restart_scheme::RestartScheme,
    restart_to_current_metric::RestartToCurrentMetric,
    restart_frequency_if_fixed::Int64,
    artificial_restart_threshold::Float64,
    sufficient_reduction_for_restart::Float64,
    necessary_reduction_for_restart::Float64,
    primal_weight_update_smoothing::Float64,
)
    # the library checks the same conditions and reports them as an error;
    # the assertions keep the reference's failure mode at construction time
    @assert restart_frequency_if_fixed > 1
    @assert 0.0 < artificial_restart_threshold <= 1.0
    @assert 0.0 < sufficient_reduction_for_restart <= necessary_reduction_for_restart <= 1.0
    @assert 0.0 <= primal_weight_update_smoothing <= 1.0
    return RestartParameters(restart_scheme, restart_to_current_metric,
                             restart_frequency_if_fixed, artificial_restart_threshold,
                             sufficient_reduction_for_restart,
                             necessary_reduction_for_restart, primal_weight_update_smoothing)
end

# The reference's eleven fields in their order, then this backend's
# extensions (include/pdlp_hip.h), which MOI's RawOptimizerAttribute can set
# by name like any other field (src/MOI_wrapper.jl:104-121).
mutable struct PdhgParameters
    l_inf_ruiz_iterations::Int
    l2_norm_rescaling::Bool
    pock_chambolle_alpha::Union{Float64,Nothing}
    primal_importance::Float64
    scale_invariant_initial_primal_weight::Bool
    verbosity::Int64
    record_iteration_stats::Bool
    termination_evaluation_frequency::Int32
    termination_criteria::TerminationCriteria
    restart_params::RestartParameters
    step_size_policy_params::Union{AdaptiveStepsizeParams,ConstantStepsizeParams}
    # ---- extensions ----
    precision::Symbol                  # :fp64 (the reference's arithmetic) or :fp32_refine
    refine_max_outer::Int32
    refine_inner_tolerance::Float64
    refine_max_scale::Float64
    device_preconditioning::Bool
    refine_inner_iteration_limit::Int32
    batch_groups::Int32
    refine_early_fallback::Int32
    refine_warm_start::Int32
    refine_continue_progress::Int32
end

# The reference's positional constructor (src/MOI_wrapper.jl:58-78 calls it
# with eleven arguments): extensions take the library's defaults.
PdhgParameters(a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11) =
    PdhgParameters(a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11,
                   :fp64, Int32(12), 1e-4, 1e6, true, Int32(0), Int32(0), Int32(0), Int32(3),
                   Int32(0))

struct SaddlePointOutput
    primal_solution::Vector{Float64}
    dual_solution::Vector{Float64}
    termination_reason::TerminationReason
    termination_string::String
    iteration_count::Int32
    iteration_stats::Vector{IterationStats}
end

# ------------------------------------------------ C mirrors (include/pdlp_hip.h)
# isbits structs: Julia lays them out as C does.

struct CLp
    n::Int64
    m::Int64
    num_eq::Int64
    nnz::Int64
    colptr::Ptr{Int64}
    rowval::Ptr{Int64}
    nzval::Ptr{Float64}
    c::Ptr{Float64}
    lb::Ptr{Float64}
    ub::Ptr{Float64}
    b::Ptr{Float64}
    c0::Float64
end

struct CTermination
    optimality_norm::Int32
    pad0::Int32
    eps_optimal_absolute::Float64
    eps_optimal_relative::Float64
    eps_primal_infeasible::Float64
    eps_dual_infeasible::Float64
    time_sec_limit::Float64
    iteration_limit::Int32
    pad1::Int32
    kkt_matrix_pass_limit::Float64
end

struct CRestart
    restart_scheme::Int32
    restart_to_current_metric::Int32
    restart_frequency_if_fixed::Int64
    artificial_restart_threshold::Float64
    sufficient_reduction_for_restart::Float64
    necessary_reduction_for_restart::Float64
    primal_weight_update_smoothing::Float64
end

struct CParams
    l_inf_ruiz_iterations::Int64
    l2_norm_rescaling::Int32
    has_pock_chambolle_alpha::Int32
    pock_chambolle_alpha::Float64
    primal_importance::Float64
    scale_invariant_initial_primal_weight::Int32
    pad0::Int32
    verbosity::Int64
    record_iteration_stats::Int32
    termination_evaluation_frequency::Int32
    termination_criteria::CTermination
    restart_params::CRestart
    step_size_policy::Int32
    pad1::Int32
    reduction_exponent::Float64
    growth_exponent::Float64
    precision::Int32
    refine_max_outer::Int32
    refine_inner_tolerance::Float64
    refine_max_scale::Float64
    device_preconditioning::Int32
    refine_inner_iteration_limit::Int32
    batch_groups::Int32
    refine_early_fallback::Int32
    debug_numerical_error_attempt::Int32
    refine_warm_start::Int32
    lane_budget::Int32
    slice_unroll::Int32
    slice_waves::Int32
    refine_continue_progress::Int32
    primal_occupancy::Int32
    dual_occupancy::Int32
end

struct CConvergence
    candidate_type::Int32
    pad::Int32
    primal_objective::Float64
    dual_objective::Float64
    corrected_dual_objective::Float64
    l_inf_primal_residual::Float64
    l2_primal_residual::Float64
    l_inf_dual_residual::Float64
    l2_dual_residual::Float64
    relative_l_inf_primal_residual::Float64
    relative_l2_primal_residual::Float64
    relative_l_inf_dual_residual::Float64
    relative_l2_dual_residual::Float64
    relative_optimality_gap::Float64
    l_inf_primal_variable::Float64
    l2_primal_variable::Float64
    l_inf_dual_variable::Float64
    l2_dual_variable::Float64
end

struct CInfeasibility
    candidate_type::Int32
    pad::Int32
    max_primal_ray_infeasibility::Float64
    primal_ray_linear_objective::Float64
    max_dual_ray_infeasibility::Float64
    dual_ray_objective::Float64
end

struct CIterationStats
    iteration_number::Int32
    cumulative_rejected_steps::Int32
    convergence_information::CConvergence
    infeasibility_information::CInfeasibility
    cumulative_kkt_matrix_passes::Float64
    cumulative_time_sec::Float64
    restart_used::Int32
    pad::Int32
    step_size::Float64
    primal_weight::Float64
    time_spent_doing_basic_algorithm::Float64
    total_attempts::Int64
end

mutable struct CResult
    x::Ptr{Float64}
    y::Ptr{Float64}
    x_current::Ptr{Float64}
    y_current::Ptr{Float64}
    termination_reason::Int32
    iteration_count::Int32
    final_stats::CIterationStats
    solve_time_sec::Float64
    preprocess_time_sec::Float64
    total_attempts::Int64
    refinement_rounds::Int32
    fp64_fallback::Int32
    CResult() = new()
end

# ------------------------------------------------------------------ conversion

function _c_params(p::PdhgParameters)
    tc = p.termination_criteria
    rp = p.restart_params
    adaptive = p.step_size_policy_params isa AdaptiveStepsizeParams
    ct = CTermination(tc.optimality_norm == L_INF ? 0 : 1, 0,
                      tc.eps_optimal_absolute, tc.eps_optimal_relative,
                      tc.eps_primal_infeasible, tc.eps_dual_infeasible,
                      tc.time_sec_limit, Int32(tc.iteration_limit), 0,
                      tc.kkt_matrix_pass_limit)
    cr = CRestart(Int32(rp.restart_scheme), Int32(rp.restart_to_current_metric),
                  rp.restart_frequency_if_fixed, rp.artificial_restart_threshold,
                  rp.sufficient_reduction_for_restart, rp.necessary_reduction_for_restart,
                  rp.primal_weight_update_smoothing)
    p.precision in (:fp64, :fp32_refine) || error("precision must be :fp64 or :fp32_refine")
    return CParams(p.l_inf_ruiz_iterations, p.l2_norm_rescaling,
                   p.pock_chambolle_alpha === nothing ? 0 : 1,
                   p.pock_chambolle_alpha === nothing ? 0.0 : p.pock_chambolle_alpha,
                   p.primal_importance, p.scale_invariant_initial_primal_weight, 0,
                   p.verbosity, p.record_iteration_stats, p.termination_evaluation_frequency,
                   ct, cr,
                   adaptive ? 0 : 1, 0,
                   adaptive ? p.step_size_policy_params.reduction_exponent : 0.3,
                   adaptive ? p.step_size_policy_params.growth_exponent : 0.6,
                   p.precision === :fp64 ? 0 : 1,   # PDLP_PRECISION_FP64 / _FP32_REFINE
                   p.refine_max_outer, p.refine_inner_tolerance, p.refine_max_scale,
                   Int32(p.device_preconditioning), p.refine_inner_iteration_limit,
                   p.batch_groups, p.refine_early_fallback, 0, p.refine_warm_start,
                   0, 0, 0,  # tiles: the library's defaults
                   p.refine_continue_progress, 0, 0)  # occupancy caps: the defaults
end

function _stats(s::CIterationStats)
    c = s.convergence_information
    i = s.infeasibility_information
    conv = ConvergenceInformation(PointType(c.candidate_type), c.primal_objective,
        c.dual_objective, c.corrected_dual_objective, c.l_inf_primal_residual,
        c.l2_primal_residual, c.l_inf_dual_residual, c.l2_dual_residual,
        c.relative_l_inf_primal_residual, c.relative_l2_primal_residual,
        c.relative_l_inf_dual_residual, c.relative_l2_dual_residual,
        c.relative_optimality_gap, c.l_inf_primal_variable, c.l2_primal_variable,
        c.l_inf_dual_variable, c.l2_dual_variable)
    infe = InfeasibilityInformation(PointType(i.candidate_type),
        i.max_primal_ray_infeasibility, i.primal_ray_linear_objective,
        i.max_dual_ray_infeasibility, i.dual_ray_objective)
    return IterationStats(s.iteration_number, [conv], [infe],
        s.cumulative_kkt_matrix_passes, s.cumulative_rejected_steps, s.cumulative_time_sec,
        RestartChoice(s.restart_used), s.step_size, s.primal_weight,
        Dict{String,Float64}("time_spent_doing_basic_algorithm" =>
                             s.time_spent_doing_basic_algorithm))
end

function _record_stats(sp::Ptr{CIterationStats}, user::Ptr{Cvoid})::Cvoid
    stats = unsafe_pointer_to_objref(user)::Vector{IterationStats}
    push!(stats, _stats(unsafe_load(sp)))
    return nothing
end

const _CTX = Ref{Ptr{Cvoid}}(C_NULL)

function _context()
    if _CTX[] == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:pdlp_ctx_create, LIBPDLP), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, out)
        rc == 0 || error("pdlp_ctx_create: " * unsafe_string(
            ccall((:pdlp_last_error, LIBPDLP), Cstring, (Ptr{Cvoid},), C_NULL)))
        _CTX[] = out[]
    end
    return _CTX[]
end

# ------------------------------------------------------------------ optimize

"""
    optimize(params::PdhgParameters, original_problem::QuadraticProgrammingProblem)

Restarted PDHG on one MI355X (src/primal_dual_hybrid_gradient_gpu.jl:457-798).
Validation errors raise like the reference's `error()` calls; the objective
matrix is ignored as on the reference's GPU path.
"""
function optimize(params::PdhgParameters, original_problem::QuadraticProgrammingProblem)
    p = original_problem
    A = p.constraint_matrix
    n, m = size(A, 2), size(A, 1)
    colptr = Vector{Int64}(A.colptr)
    rowval = Vector{Int64}(A.rowval)
    nzval = Vector{Float64}(A.nzval)
    c, lb, ub, b = p.objective_vector, p.variable_lower_bound, p.variable_upper_bound,
                   p.right_hand_side
    x = zeros(Float64, n)
    y = zeros(Float64, m)
    stats = IterationStats[]
    res = CResult()
    cb = @cfunction(_record_stats, Cvoid, (Ptr{CIterationStats}, Ptr{Cvoid}))
    GC.@preserve colptr rowval nzval c lb ub b x y stats begin
        lp = CLp(n, m, p.num_equalities, length(nzval), pointer(colptr), pointer(rowval),
                 pointer(nzval), pointer(c), pointer(lb), pointer(ub), pointer(b),
                 p.objective_constant)
        res.x = pointer(x)
        res.y = pointer(y)
        res.x_current = C_NULL
        res.y_current = C_NULL
        ctx = _context()
        rc = ccall((:pdlp_solve, LIBPDLP), Cint,
                   (Ptr{Cvoid}, Ref{CLp}, Ref{CParams}, Ref{CResult}, Ptr{Cvoid}, Any),
                   ctx, lp, _c_params(params), res, cb, stats)
        if rc != 0
            msg = unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring, (Ptr{Cvoid},), ctx))
            rc > 0 ? error(msg) : throw(ErrorException("HIP failure: " * msg))
        end
    end
    final = _stats(res.final_stats)
    if isempty(stats) || stats[end].iteration_number != final.iteration_number
        push!(stats, final)
    else
        stats[end] = final
    end
    reason = TerminationReason(res.termination_reason)
    # termination_reason_to_string: src/termination.jl:248-251 (the string
    # MOI's RawStatusString prints, src/MOI_wrapper.jl:216-221)
    return SaddlePointOutput(x, y, reason, termination_reason_to_string(reason),
                             res.iteration_count, stats)
end

function _c_lp(p::QuadraticProgrammingProblem, keep::Vector{Any})
    A = p.constraint_matrix
    colptr = Vector{Int64}(A.colptr)
    rowval = Vector{Int64}(A.rowval)
    nzval = Vector{Float64}(A.nzval)
    c, lb, ub, b = p.objective_vector, p.variable_lower_bound, p.variable_upper_bound,
                   p.right_hand_side
    append!(keep, (colptr, rowval, nzval, c, lb, ub, b))
    return CLp(size(A, 2), size(A, 1), p.num_equalities, length(nzval), pointer(colptr),
               pointer(rowval), pointer(nzval), pointer(c), pointer(lb), pointer(ub),
               pointer(b), p.objective_constant)
end

function _output(x, y, res::CResult)
    reason = TerminationReason(res.termination_reason)
    return SaddlePointOutput(x, y, reason, termination_reason_to_string(reason),
                             res.iteration_count, [_stats(res.final_stats)])
end

"""
    batch_optimize(params::PdhgParameters, problems::Vector{QuadraticProgrammingProblem})

Solves independent LPs together on the device (pdlp_batch_solve): packed
block-diagonally into the same kernel launches, each with its own step
sizes, restarts and termination, each result equal to its own `optimize`
(without the iteration history: each output holds the final statistics).
`params.precision = :fp32_refine` runs the FP32 + FP64 refinement form.
"""
function batch_optimize(params::PdhgParameters, problems::Vector{QuadraticProgrammingProblem})
    k = length(problems)
    keep = Any[]
    xs = [zeros(Float64, size(p.constraint_matrix, 2)) for p in problems]
    ys = [zeros(Float64, size(p.constraint_matrix, 1)) for p in problems]
    lps = Vector{CLp}(undef, k)
    results = [CResult() for _ in 1:k]
    GC.@preserve keep xs ys begin
        for i in 1:k
            lps[i] = _c_lp(problems[i], keep)
            results[i].x = pointer(xs[i])
            results[i].y = pointer(ys[i])
            results[i].x_current = C_NULL
            results[i].y_current = C_NULL
        end
        # pdlp_result is read and written in place: pass a contiguous array of
        # the isbits images and copy back
        flat = [_result_image(r) for r in results]
        ctx = _context()
        rc = ccall((:pdlp_batch_solve, LIBPDLP), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{CLp}, Ref{CParams}, Ptr{CResultImage}),
                   ctx, Int32(k), lps, _c_params(params), flat)
        if rc != 0
            msg = unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring, (Ptr{Cvoid},), ctx))
            rc > 0 ? error(msg) : throw(ErrorException("HIP failure: " * msg))
        end
        return [_output(xs[i], ys[i], _from_image(flat[i])) for i in 1:k]
    end
end

# isbits image of pdlp_result, for arrays of results
struct CResultImage
    x::Ptr{Float64}
    y::Ptr{Float64}
    x_current::Ptr{Float64}
    y_current::Ptr{Float64}
    termination_reason::Int32
    iteration_count::Int32
    final_stats::CIterationStats
    solve_time_sec::Float64
    preprocess_time_sec::Float64
    total_attempts::Int64
    refinement_rounds::Int32
    fp64_fallback::Int32
end

_result_image(r::CResult) = CResultImage(r.x, r.y, r.x_current, r.y_current, 0, 0,
    _zero_stats(), 0.0, 0.0, 0, 0, 0)
_zero_stats() = reinterpret(CIterationStats, zeros(UInt8, sizeof(CIterationStats)))[1]

function _from_image(im::CResultImage)
    r = CResult()
    r.x, r.y, r.x_current, r.y_current = im.x, im.y, im.x_current, im.y_current
    r.termination_reason = im.termination_reason
    r.iteration_count = im.iteration_count
    r.final_stats = im.final_stats
    r.solve_time_sec = im.solve_time_sec
    r.preprocess_time_sec = im.preprocess_time_sec
    r.total_attempts = im.total_attempts
    r.refinement_rounds = im.refinement_rounds
    r.fp64_fallback = im.fp64_fallback
    return r
end

# ------------------------------------------------------------ request loop
# Continuous batching (pdlp_serve_*): requests are admitted by HBM footprint
# as slots free up, run in shared launches, and retire as they converge.

mutable struct HipServer
    handle::Ptr{Cvoid}
    sizes::Dict{Int64,Tuple{Int,Int}}
end

function HipServer(params::PdhgParameters; hbm_budget_bytes::Float64 = 200e9,
                   max_slots::Integer = 0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    ctx = _context()
    rc = ccall((:pdlp_serve_open, LIBPDLP), Cint,
               (Ptr{Cvoid}, Ref{CParams}, Float64, Int32, Ref{Ptr{Cvoid}}),
               ctx, _c_params(params), hbm_budget_bytes, Int32(max_slots), out)
    rc == 0 || error(unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring,
                                         (Ptr{Cvoid},), ctx)))
    return HipServer(out[], Dict{Int64,Tuple{Int,Int}}())
end

"""
Queues an LP with its own tolerance (and optionally its own time limit,
counted from submission, and iteration limit); returns the request id.
"""
function submit!(s::HipServer, p::QuadraticProgrammingProblem, tolerance::Float64;
                 time_sec_limit::Float64 = Inf, iteration_limit::Integer = -1)
    keep = Any[]
    id = Ref{Int64}(0)
    GC.@preserve keep begin
        lp = _c_lp(p, keep)
        rc = ccall((:pdlp_serve_submit_limits, LIBPDLP), Cint,
                   (Ptr{Cvoid}, Ref{CLp}, Float64, Float64, Int32, Ref{Int64}),
                   s.handle, lp, tolerance, time_sec_limit, Int32(iteration_limit), id)
        rc == 0 || error(unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring,
                                             (Ptr{Cvoid},), _context())))
    end
    s.sizes[id[]] = (size(p.constraint_matrix, 2), size(p.constraint_matrix, 1))
    return id[]
end

"Advances the device work; returns the ids finished since the last poll."
function poll!(s::HipServer)
    ids = zeros(Int64, 256)
    got = Ref{Int32}(0)
    rc = ccall((:pdlp_serve_poll, LIBPDLP), Cint,
               (Ptr{Cvoid}, Ptr{Int64}, Int32, Ref{Int32}), s.handle, ids, Int32(256), got)
    rc == 0 || error(unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring,
                                         (Ptr{Cvoid},), _context())))
    return ids[1:got[]]
end

"Cancels a queued or running request (it finishes with TERMINATION_REASON_OTHER)."
function cancel!(s::HipServer, id::Int64)
    rc = ccall((:pdlp_serve_cancel, LIBPDLP), Cint, (Ptr{Cvoid}, Int64), s.handle, id)
    rc == 0 || error(unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring,
                                         (Ptr{Cvoid},), _context())))
    return
end

"The SaddlePointOutput of a finished request."
function result!(s::HipServer, id::Int64)
    n, m = s.sizes[id]
    x, y = zeros(Float64, n), zeros(Float64, m)
    res = CResult()
    GC.@preserve x y begin
        res.x, res.y = pointer(x), pointer(y)
        res.x_current, res.y_current = C_NULL, C_NULL
        rc = ccall((:pdlp_serve_result, LIBPDLP), Cint, (Ptr{Cvoid}, Int64, Ref{CResult}),
                   s.handle, id, res)
        rc == 0 || error(unsafe_string(ccall((:pdlp_last_error, LIBPDLP), Cstring,
                                             (Ptr{Cvoid},), _context())))
    end
    delete!(s.sizes, id)
    return _output(x, y, res)
end

function Base.close(s::HipServer)
    s.handle == C_NULL && return
    ccall((:pdlp_serve_close, LIBPDLP), Cint, (Ptr{Cvoid},), s.handle)
    s.handle = C_NULL
    return
end

# ---------------------------------------------------------------- MPS input
# qps_reader_to_standard_form (src/quadratic_programming_io.jl:147-186) over
# pdlp_read_mps: the file is parsed and put in standard form natively (the
# QPSReader and GZip dependencies go away).  Include this after
# quadratic_programming_io.jl to replace its definition, or in its place.

struct CMpsProblem
    lp::CLp
    num_file_variables::Int64
    num_slack_variables::Int64
    quadratic_entries::Int64
    owner::Ptr{Cvoid}
end

function qps_reader_to_standard_form(filename::String; fixed_format::Bool = false)
    out = Ref{Ptr{CMpsProblem}}(C_NULL)
    err = zeros(UInt8, 512)
    rc = ccall((:pdlp_read_mps, LIBPDLP), Cint,
               (Cstring, Int32, Ref{Ptr{CMpsProblem}}, Ptr{UInt8}, Int64),
               filename, Int32(fixed_format), out, err, length(err))
    rc == 0 || error(unsafe_string(pointer(err)))
    try
        p = unsafe_load(out[])
        lp = p.lp
        n, m, nnz = Int(lp.n), Int(lp.m), Int(lp.nnz)
        p.quadratic_entries == 0 || error("quadratic objectives are not supported by the GPU path")
        vec(ptr, len) = copy(unsafe_wrap(Vector{eltype(ptr)}, ptr, len))
        A = SparseMatrixCSC(m, n, vec(lp.colptr, n + 1), vec(lp.rowval, nnz), vec(lp.nzval, nnz))
        return QuadraticProgrammingProblem(vec(lp.lb, n), vec(lp.ub, n), spzeros(n, n),
                                           vec(lp.c, n), lp.c0, A, vec(lp.b, m), Int(lp.num_eq))
    finally
        ccall((:pdlp_mps_free, LIBPDLP), Cvoid, (Ptr{CMpsProblem},), out[])
    end
end
