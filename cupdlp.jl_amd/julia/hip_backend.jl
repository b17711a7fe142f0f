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
end

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
    iteration_limit::Int32
    eps_optimal_absolute::Float64
    eps_optimal_relative::Float64
    eps_primal_infeasible::Float64
    eps_dual_infeasible::Float64
    time_sec_limit::Float64
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
    record_iteration_stats::Int32
    verbosity::Int64
    termination_evaluation_frequency::Int32
    step_size_policy::Int32
    reduction_exponent::Float64
    growth_exponent::Float64
    termination_criteria::CTermination
    restart_params::CRestart
    precision::Int32
    refine_max_outer::Int32
    refine_inner_tolerance::Float64
    refine_max_scale::Float64
    device_preconditioning::Int32
    refine_inner_iteration_limit::Int32
    batch_groups::Int32
    refine_early_fallback::Int32
    reserved::NTuple{4,Int32}
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
    ct = CTermination(tc.optimality_norm == L_INF ? 0 : 1, Int32(tc.iteration_limit),
                      tc.eps_optimal_absolute, tc.eps_optimal_relative,
                      tc.eps_primal_infeasible, tc.eps_dual_infeasible,
                      tc.time_sec_limit, tc.kkt_matrix_pass_limit)
    cr = CRestart(Int32(rp.restart_scheme), Int32(rp.restart_to_current_metric),
                  rp.restart_frequency_if_fixed, rp.artificial_restart_threshold,
                  rp.sufficient_reduction_for_restart, rp.necessary_reduction_for_restart,
                  rp.primal_weight_update_smoothing)
    return CParams(p.l_inf_ruiz_iterations, p.l2_norm_rescaling,
                   p.pock_chambolle_alpha === nothing ? 0 : 1,
                   p.pock_chambolle_alpha === nothing ? 0.0 : p.pock_chambolle_alpha,
                   p.primal_importance, p.scale_invariant_initial_primal_weight,
                   p.record_iteration_stats, p.verbosity, p.termination_evaluation_frequency,
                   adaptive ? 0 : 1,
                   adaptive ? p.step_size_policy_params.reduction_exponent : 0.3,
                   adaptive ? p.step_size_policy_params.growth_exponent : 0.6,
                   ct, cr,
                   0,        # PDLP_PRECISION_FP64: the reference's arithmetic
                   12, 1e-4, 1e6, 1, 0, 0, 0, ntuple(_ -> Int32(0), 4))
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
    return SaddlePointOutput(x, y, reason, "", res.iteration_count, stats)
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
