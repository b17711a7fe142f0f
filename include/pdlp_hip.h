/*
 * pdlp_hip.h — C ABI of the MI355X restarted-PDHG LP engine (libpdlp_hip.so).
 *
 * This is the drop-in boundary that replaces the body of cuPDLP.jl's
 *   optimize(params::PdhgParameters, original_problem::QuadraticProgrammingProblem)
 *       ::SaddlePointOutput            (src/primal_dual_hybrid_gradient_gpu.jl:457-798)
 * A Julia `cuPDLP` module keeps its public types and `cuPDLP.Optimizer`
 * (src/MOI_wrapper.jl) and `ccall`s the functions below.  Only plain C types
 * cross the boundary: no C++ or torch types, no CUDA/HIP types.
 *
 * Conventions
 *  - The LP is given exactly as Julia's SparseMatrixCSC{Float64,Int64} holds
 *    it: 1-based CSC (colptr[n+1], rowval[nnz], nzval[nnz]).  Host arrays are
 *    borrowed for the duration of a call only and never retained.
 *  - Rows 1..num_eq are equalities, rows num_eq+1..m are ">=" rows
 *    (src/quadratic_programming.jl:6-22).
 *  - Enum integer values equal the reference's Julia @enum values (first = 0).
 *  - Return codes: 0 ok; > 0 invalid input (message in pdlp_last_error);
 *    < 0 HIP / runtime failure (message in pdlp_last_error).
 */
#ifndef PDLP_HIP_H
#define PDLP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDLP_ABI_VERSION 2

/* TerminationReason — src/solve_log.jl:323-334 */
enum {
  PDLP_TERMINATION_REASON_UNSPECIFIED = 0,
  PDLP_TERMINATION_REASON_OPTIMAL = 1,
  PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE = 2,
  PDLP_TERMINATION_REASON_DUAL_INFEASIBLE = 3,
  PDLP_TERMINATION_REASON_TIME_LIMIT = 4,
  PDLP_TERMINATION_REASON_ITERATION_LIMIT = 5,
  PDLP_TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT = 6,
  PDLP_TERMINATION_REASON_NUMERICAL_ERROR = 7,
  PDLP_TERMINATION_REASON_INVALID_PROBLEM = 8,
  PDLP_TERMINATION_REASON_OTHER = 9
};

/* RestartChoice — src/solve_log.jl:19-24 */
enum {
  PDLP_RESTART_CHOICE_UNSPECIFIED = 0,
  PDLP_RESTART_CHOICE_NO_RESTART = 1,
  PDLP_RESTART_CHOICE_WEIGHTED_AVERAGE_RESET = 2,
  PDLP_RESTART_CHOICE_RESTART_TO_AVERAGE = 3
};

/* PointType — src/solve_log.jl:39-45 */
enum {
  PDLP_POINT_TYPE_UNSPECIFIED = 0,
  PDLP_POINT_TYPE_CURRENT_ITERATE = 1,
  PDLP_POINT_TYPE_ITERATE_DIFFERENCE = 2,
  PDLP_POINT_TYPE_AVERAGE_ITERATE = 3,
  PDLP_POINT_TYPE_NONE = 4
};

/* RestartScheme, RestartToCurrentMetric — src/saddle_point_gpu.jl:343,350 */
enum { PDLP_NO_RESTARTS = 0, PDLP_FIXED_FREQUENCY = 1, PDLP_ADAPTIVE_KKT = 2 };
enum { PDLP_NO_RESTART_TO_CURRENT = 0, PDLP_KKT_GREEDY = 1 };

/* OptimalityNorm — src/termination.jl:6 */
enum { PDLP_L_INF = 0, PDLP_L2 = 1 };

/* step_size_policy_params variant — src/primal_dual_hybrid_gradient_gpu.jl:2-7 */
enum { PDLP_STEP_ADAPTIVE = 0, PDLP_STEP_CONSTANT = 1 };

/* Arithmetic of the bulk iteration (extension, no reference counterpart). */
enum {
  PDLP_PRECISION_FP64 = 0,        /* the reference's arithmetic */
  PDLP_PRECISION_FP32_REFINE = 1  /* FP32 matrix/iterates + FP64 iterative refinement */
};

/* Borrowed view of an LP in the reference's host layout
 * (QuadraticProgrammingProblem, src/quadratic_programming.jl:25-67, with the
 * objective matrix absent: the reference's GPU path ignores it,
 * src/cpu_to_gpu.jl:1-14). */
typedef struct pdlp_lp {
  int64_t n;            /* number of variables */
  int64_t m;            /* number of constraint rows */
  int64_t num_eq;       /* rows [0, num_eq) are equalities */
  int64_t nnz;          /* colptr[n] - 1 */
  const int64_t *colptr;  /* n+1, 1-based */
  const int64_t *rowval;  /* nnz, 1-based row indices, sorted within a column */
  const double *nzval;    /* nnz */
  const double *c;        /* objective_vector, n */
  const double *lb;       /* variable_lower_bound, n (may hold -Inf) */
  const double *ub;       /* variable_upper_bound, n (may hold +Inf) */
  const double *b;        /* right_hand_side, m */
  double c0;              /* objective_constant */
} pdlp_lp;

/* TerminationCriteria — src/termination.jl:9-77 (same field order; pad_
 * fields are C alignment only). */
typedef struct pdlp_termination_criteria {
  int32_t optimality_norm;          /* PDLP_L_INF / PDLP_L2 */
  int32_t pad0_;
  double eps_optimal_absolute;
  double eps_optimal_relative;
  double eps_primal_infeasible;
  double eps_dual_infeasible;
  double time_sec_limit;
  int32_t iteration_limit;          /* Int32 in the reference */
  int32_t pad1_;
  double kkt_matrix_pass_limit;
} pdlp_termination_criteria;

/* RestartParameters — src/saddle_point_gpu.jl:353-383 (same field order). */
typedef struct pdlp_restart_params {
  int32_t restart_scheme;
  int32_t restart_to_current_metric;
  int64_t restart_frequency_if_fixed;
  double artificial_restart_threshold;
  double sufficient_reduction_for_restart;
  double necessary_reduction_for_restart;
  double primal_weight_update_smoothing;
} pdlp_restart_params;

/* PdhgParameters — src/primal_dual_hybrid_gradient_gpu.jl:9-24, field for
 * field in the same order (pad_ fields are C alignment only):
 *   pock_chambolle_alpha::Union{Float64,Nothing}  -> has_pock_chambolle_alpha
 *                                                    + pock_chambolle_alpha
 *   step_size_policy_params::Union{AdaptiveStepsizeParams,
 *                                  ConstantStepsizeParams}
 *                                                 -> step_size_policy (the
 *                                                    variant) + the fields of
 *                                                    AdaptiveStepsizeParams (:1-4)
 * followed by extensions that have no reference counterpart. */
typedef struct pdlp_params {
  int64_t l_inf_ruiz_iterations;
  int32_t l2_norm_rescaling;
  int32_t has_pock_chambolle_alpha;  /* 0 encodes `nothing` */
  double pock_chambolle_alpha;
  double primal_importance;
  int32_t scale_invariant_initial_primal_weight;
  int32_t pad0_;
  int64_t verbosity;
  int32_t record_iteration_stats;
  int32_t termination_evaluation_frequency;
  pdlp_termination_criteria termination_criteria;
  pdlp_restart_params restart_params;
  int32_t step_size_policy;          /* PDLP_STEP_ADAPTIVE / PDLP_STEP_CONSTANT */
  int32_t pad1_;
  double reduction_exponent;         /* AdaptiveStepsizeParams */
  double growth_exponent;
  /* ---- extensions ---- */
  int32_t precision;                 /* PDLP_PRECISION_* */
  int32_t refine_max_outer;          /* FP32_REFINE: max refinement rounds */
  double refine_inner_tolerance;     /* FP32_REFINE: tolerance of the first FP32 solve and
                                        the tightest asked of a correction solve, which
                                        asks for max(this, min(0.1, 0.1 / e)), e = the
                                        refined point's error over the outer tolerance */
  double refine_max_scale;           /* FP32_REFINE: cap of the per-round scale-up */
  int32_t device_preconditioning;    /* 1: Ruiz/Pock-Chambolle on the GPU */
  int32_t refine_inner_iteration_limit; /* FP32_REFINE: iteration cap of each correction
                                           solve; 0 = max(10000, 20 x the first solve's) */
  int32_t batch_groups;              /* batched solves: size groups run side by side, each
                                        with its own launches and streams; 0 = automatic
                                        (3 from 24 instances up), 1 = one group.  Servers:
                                        size classes by nonzeros (< 2^17, < 2^20, < 2^23,
                                        above), each with its own engine and stream;
                                        0 = 3, at most 4 */
  int32_t refine_early_fallback;     /* FP32_REFINE, pdlp_batch_solve / pdlp_batch_run:
                                        iterations into a correction solve after which
                                        the instance's plain FP64 solve starts beside it,
                                        used only if the refinement then stalls (results
                                        are those of starting it at the stall); <= 0 =
                                        never (the default: the extra FP64 work costs the
                                        512-LP batch more than it saves) */
  int32_t debug_numerical_error_attempt; /* fault injection for tests: > 0 makes the
                                            step attempt with this total_number_iterations
                                            see movement = 0, the reference's numerical
                                            error branch (src/primal_dual_hybrid_gradient_gpu.jl:
                                            384-390) when its interaction is > 0; 0 = off */
  int32_t refine_warm_start;         /* FP32_REFINE: a correction solve starts from the
                                        previous solve's final primal weight (bit 1,
                                        rescaled by the correction's scaling) and step
                                        size (bit 2) instead of fresh ones; 0 = fresh;
                                        default 3 (fewer stalled refinements on the
                                        512-LP batch: 16 vs 21) */
  /* Tile parameters of the step kernels (pdlp_tune picks them by timed
   * trials; 0 = the default): */
  int32_t lane_budget;               /* most entries of a row one lane of a slice sums
                                        (the lanes per row follow; 16..1024, default 64);
                                        changes the slicing, hence the order of a row's
                                        compensated sum */
  int32_t slice_unroll;              /* loads in flight per lane: 4, 8 or 16 for both
                                        slice kernels; 0 (default): 8, and 16 for the
                                        FP64 dual kernel */
  int32_t slice_waves;               /* waves per work-group of the slice kernels: 1, 2
                                        or 4 (default 4); unroll and waves change no
                                        result */
  int32_t refine_continue_progress; /* FP32_REFINE: a correction solve stopped by its
                                        iteration cap (after the first solve) lets the
                                        refinement go on when it cut the refined point's
                                        error at least this many times; otherwise, or
                                        when 0 (the default), the refinement counts as
                                        stalled and the FP64 fallback takes over.  On the
                                        512-LP batch at 1e-8 a second capped round
                                        rarely finishes (factor 10: 14 instead of 16
                                        fallbacks, 42.9 vs 47.9 solves/s) */
  int32_t primal_occupancy;          /* most waves per SIMD of the primal / dual slice
                                        kernels, held by reserving LDS per work-group at
                                        launch (fewer waves in flight thrash the caches
                                        less with their gathers); 0 = the default (5
                                        primal, 3 dual), 8 or more = as many as the
                                        registers allow.  Changes no result */
  int32_t dual_occupancy;
} pdlp_params;

/* ConvergenceInformation — src/solve_log.jl:51-155 (same field order). */
typedef struct pdlp_convergence_information {
  int32_t candidate_type;
  int32_t pad_;
  double primal_objective;
  double dual_objective;
  double corrected_dual_objective;
  double l_inf_primal_residual;
  double l2_primal_residual;
  double l_inf_dual_residual;
  double l2_dual_residual;
  double relative_l_inf_primal_residual;
  double relative_l2_primal_residual;
  double relative_l_inf_dual_residual;
  double relative_l2_dual_residual;
  double relative_optimality_gap;
  double l_inf_primal_variable;
  double l2_primal_variable;
  double l_inf_dual_variable;
  double l2_dual_variable;
} pdlp_convergence_information;

/* InfeasibilityInformation — src/solve_log.jl:161-212 (same field order). */
typedef struct pdlp_infeasibility_information {
  int32_t candidate_type;
  int32_t pad_;
  double max_primal_ray_infeasibility;
  double primal_ray_linear_objective;
  double max_dual_ray_infeasibility;
  double dual_ray_objective;
} pdlp_infeasibility_information;

/* IterationStats — src/solve_log.jl:219-302.  Exactly one convergence and
 * one infeasibility record, as the reference produces
 * (src/iteration_stats_utils_gpu.jl:405-429). */
typedef struct pdlp_iteration_stats {
  int32_t iteration_number;
  int32_t cumulative_rejected_steps;
  pdlp_convergence_information convergence_information;
  pdlp_infeasibility_information infeasibility_information;
  double cumulative_kkt_matrix_passes;
  double cumulative_time_sec;
  int32_t restart_used;
  int32_t pad_;
  double step_size;
  double primal_weight;
  double time_spent_doing_basic_algorithm;  /* method_specific_stats entry */
  int64_t total_attempts;                   /* extension: step attempts so far */
} pdlp_iteration_stats;

/* Output of one solve — SaddlePointOutput (src/saddle_point_gpu.jl:3-34)
 * with caller-allocated vectors.  x (n) and y (m) receive the unscaled
 * average iterate the reference returns (src/saddle_point_gpu.jl:40-61). */
typedef struct pdlp_result {
  double *x;                 /* caller-allocated, n */
  double *y;                 /* caller-allocated, m */
  double *x_current;         /* optional (NULL): unscaled current iterate, n */
  double *y_current;         /* optional (NULL): unscaled current iterate, m */
  int32_t termination_reason;
  int32_t iteration_count;
  pdlp_iteration_stats final_stats;
  double solve_time_sec;         /* wall time of the whole call (pdlp_batch_run:
                                    from the start of the run) */
  double preprocess_time_sec;    /* validation + layout + upload + rescaling */
  int64_t total_attempts;        /* total_number_iterations of the reference */
  int32_t refinement_rounds;     /* FP32_REFINE: outer rounds performed */
  int32_t fp64_fallback;         /* FP32_REFINE: 1 if refinement stalled and the
                                    LP was solved again in FP64 (pdlp_solve,
                                    pdlp_batch_solve, the serving loop; x, y,
                                    the statistics and the reason are then the
                                    FP64 solve's, except that a solve stopped
                                    by the caller's limit further from optimal
                                    than the refined point returns that point
                                    and its statistics) */
} pdlp_result;

/* Receives each recorded IterationStats (replaces Vector{IterationStats}). */
typedef void (*pdlp_stats_cb)(const pdlp_iteration_stats *stats, void *user);

/* Opaque per-device context: one HIP stream and its device buffers.  One
 * solve at a time per context; separate contexts may be used from separate
 * host threads. */
typedef struct pdlp_ctx pdlp_ctx;

/* Creates a context on HIP device `device`. */
int pdlp_ctx_create(int device, pdlp_ctx **out);
void pdlp_ctx_destroy(pdlp_ctx *ctx);

/* Last error message of this context (or of the process if ctx is NULL). */
const char *pdlp_last_error(const pdlp_ctx *ctx);

/* Fills `p` with the defaults of cuPDLP.Optimizer()
 * (src/MOI_wrapper.jl:52-80, construct_termination_criteria
 * src/termination.jl:79-99). */
void pdlp_default_params(pdlp_params *p);

/* Replaces optimize(params, problem) — src/primal_dual_hybrid_gradient_gpu.jl:457-798. */
int pdlp_solve(pdlp_ctx *ctx, const pdlp_lp *lp, const pdlp_params *params,
               pdlp_result *result, pdlp_stats_cb cb, void *user);

/* pdlp_solve that allocates nothing (no counterpart in the reference): for
 * callers that cannot call hipMalloc -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  Same arguments, result
 * and error behaviour as pdlp_solve (bitwise the same result), except:
 *  - every device buffer comes from one fixed region per device, a
 *    file-scope array of 128 MiB (room for an FP64 LP of about 1.5 million
 *    nonzeros) whose address pdlp_ctx_create takes; no hipMalloc, hipFree,
 *    hipHostMalloc or event creation, host staging in ordinary memory;
 *  - an LP that does not fit the region is refused with code 7;
 *  - PDLP_PRECISION_FP64 only (code 2 otherwise);
 *  - the region is shared by every call on that device and keeps the last
 *    call's contents: calls that overlap on one device, from any context or
 *    thread, race and must not be made. */
int pdlp_solve_noalloc(pdlp_ctx *ctx, const pdlp_lp *lp, const pdlp_params *params,
                       pdlp_result *result, pdlp_stats_cb cb, void *user);

/* Solves `count` independent LPs together, packed block-diagonally into the
 * same kernel launches with per-instance step sizes, restarts and
 * termination (extension; each instance follows optimize() exactly). */
int pdlp_batch_solve(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps,
                     const pdlp_params *params, pdlp_result *results);

/* pdlp_batch_solve in two halves (extension), so that the LPs can be made
 * resident in HBM before the solve is timed or scheduled.  optimize() copies
 * the problem to the GPU (src/primal_dual_hybrid_gradient_gpu.jl:488-490)
 * before its solve clock starts (:602); here the rescaling, done on the
 * device, belongs to the run.
 * pdlp_batch_load validates the LPs, builds their device layout on the host
 * and copies them to the device; the caller's arrays may be freed on return.
 * pdlp_batch_run rescales on the device, solves and writes `results` exactly
 * as pdlp_batch_solve would (bitwise); solve_time_sec counts from the start
 * of the run.  A loaded batch is solved once; free it with pdlp_batch_free
 * before its context is destroyed.  Runs on one context do not overlap. */
typedef struct pdlp_batch pdlp_batch;
int pdlp_batch_load(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps,
                    const pdlp_params *params, pdlp_batch **out);
int pdlp_batch_run(pdlp_batch *batch, pdlp_result *results);
void pdlp_batch_free(pdlp_batch *batch);

/* An LP read from an MPS / QPS file in the reference's standard form --
 * replaces qps_reader_to_standard_form (src/quadratic_programming_io.jl:147-186)
 * with transform_to_standard_form (:33-82) and two_sided_rows_to_slacks
 * (:91-127): equality rows first, ">=" rows after, "<=" rows negated, a slack
 * column per two-sided row (after the file's columns).  Host-only (no
 * device).  `lp` points into memory the library owns until pdlp_mps_free. */
typedef struct pdlp_mps_problem {
  pdlp_lp lp;
  int64_t num_file_variables;   /* columns of the file (slacks follow) */
  int64_t num_slack_variables;
  int64_t quadratic_entries;    /* nonzero QUADOBJ / QMATRIX entries read */
  void *owner;                  /* library-owned storage */
} pdlp_mps_problem;
/* Reads `path` (free MPS unless fixed_format; gzip-compressed files are read
 * as such).  Returns 0, or > 0 with a message in err (err_len bytes):
 * 1 bad argument, 2 unreadable / malformed file, 3 a row with no finite
 * bound (the reference's error). */
int pdlp_read_mps(const char *path, int32_t fixed_format, pdlp_mps_problem **out, char *err,
                  int64_t err_len);
void pdlp_mps_free(pdlp_mps_problem *p);

/* Continuous-batching request loop (extension). */
typedef struct pdlp_server pdlp_server;
int pdlp_serve_open(pdlp_ctx *ctx, const pdlp_params *params,
                    double hbm_budget_bytes, int32_t max_slots,
                    pdlp_server **out);
/* Copies the LP to the device queue; returns a request id >= 0 in *id. */
int pdlp_serve_submit(pdlp_server *srv, const pdlp_lp *lp, double tolerance,
                      int64_t *id);
/* pdlp_serve_submit with limits of this request's own, applied as the
 * reference applies TerminationCriteria's (src/termination.jl:236-245) and
 * on top of the server's: time_sec_limit counts from the request's
 * submission, so time spent queued counts against it (<= 0 or +Inf: none);
 * iteration_limit < 0: none.  A request that reaches one finishes with
 * TIME_LIMIT / ITERATION_LIMIT and its last evaluated point, like a solve. */
int pdlp_serve_submit_limits(pdlp_server *srv, const pdlp_lp *lp, double tolerance,
                             double time_sec_limit, int32_t iteration_limit, int64_t *id);
/* Cancels a queued or running request.  It finishes at once (its id is
 * reported by the next poll) with termination reason OTHER, zero x and y
 * and the counts of the work done so far; its device memory is released.
 * Returns 6 for an id that is unknown or already finished. */
int pdlp_serve_cancel(pdlp_server *srv, int64_t id);
/* Advances the device work; returns the number of requests finished since
 * the last poll in *finished and copies their ids to ids[0..*finished). */
int pdlp_serve_poll(pdlp_server *srv, int64_t *ids, int32_t max_ids,
                    int32_t *finished);
/* Copies the result of a finished request (x: n, y: m caller-allocated). */
int pdlp_serve_result(pdlp_server *srv, int64_t id, pdlp_result *result);
int pdlp_serve_close(pdlp_server *srv);

/* Timing of the two iteration kernels (extension, for benchmarks): when
 * enabled, every k_primal / k_dual launch of pdlp_solve / pdlp_batch_solve is
 * bracketed by HIP events on the context's stream.  Bytes are the algorithmic
 * bytes of each launch, summed over the instances active in it:
 * N(s_A+4) + 4(rows+1) + s_v*(gathered length) + 9*s_v*rows, with
 * s_A = s_v = 8 (FP64) or 4 (FP32_REFINE). */
typedef struct pdlp_kernel_profile {
  int64_t primal_launches;
  int64_t dual_launches;
  double primal_ms;     /* summed event-to-event durations */
  double dual_ms;
  double primal_bytes;  /* summed algorithmic bytes */
  double dual_bytes;
  /* the same for the FP64 engines that finish stalled FP32 + refinement
     instances (pdlp_batch_solve / pdlp_batch_run; s_A = s_v = 8) */
  int64_t fallback_primal_launches;
  int64_t fallback_dual_launches;
  double fallback_primal_ms;
  double fallback_dual_ms;
  double fallback_primal_bytes;
  double fallback_dual_bytes;
} pdlp_kernel_profile;
/* enable: 0 off; 1 the main engines' launches timed, the fallback engines'
 * launches and bytes counted without events (their fallback_*_ms stay 0);
 * 2 the fallback engines' launches timed as well (events on every FP64
 * fallback launch: slower). */
int pdlp_set_profiling(pdlp_ctx *ctx, int32_t enable);
/* Copies the accumulated profile and resets it. */
int pdlp_get_profile(pdlp_ctx *ctx, pdlp_kernel_profile *out);

/* Tile parameters of the step kernels chosen by timed trials on the device
 * (extension; the reference hard-wires its choices: cuSPARSE's CSR_ALG2 at
 * src/primal_dual_hybrid_gradient_gpu.jl:225, 307 and ThreadPerBlock = 128
 * at src/cuPDLP.jl:34). */
typedef struct pdlp_tuning {
  int32_t lane_budget;              /* the fields to copy into pdlp_params */
  int32_t slice_unroll;
  int32_t slice_waves;
  int32_t candidates;               /* configurations timed (0: taken from the cache) */
  double kernel_us_per_iteration;   /* k_primal + k_dual per launch pair, the choice */
  double default_us_per_iteration;  /* the same with the defaults (64, 0, 4) */
  double seconds;                   /* wall time of the trials */
  int32_t primal_occupancy;         /* the fields to copy into pdlp_params */
  int32_t dual_occupancy;
} pdlp_tuning;
/* Solves the LPs (batched as pdlp_batch_solve would, with params) for
 * trial_iterations iterations per candidate and times their step kernels
 * with HIP events: lane budget 32 / 64 / 128, then 2 or 4 waves per
 * work-group, then the default loads in flight or 4 (FP64: also 8 for the
 * dual kernel as for the primal one), each step keeping a candidate
 * only if it is at least 3 % faster; then, with the best of those, at most
 * 3, 4, 5, 6 or 8 waves per SIMD, taken for the primal and the dual kernel
 * separately (each the fastest of its own launches).  Writes the fastest to *out.  The
 * choice is cached on the context per LP shape (count, total n, m and
 * nonzeros, precision, size groups): a later call for the same shape
 * returns it at once.  Unroll and waves change no result; the lane budget
 * changes the order of each row's compensated sum. */
int pdlp_tune(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps, const pdlp_params *params,
              int32_t trial_iterations, pdlp_tuning *out);

/* Library identification. */
int pdlp_abi_version(void);

/* Writes sizeof() of pdlp_lp, pdlp_termination_criteria, pdlp_restart_params,
 * pdlp_params, pdlp_convergence_information, pdlp_infeasibility_information,
 * pdlp_iteration_stats and pdlp_result (in that order) to out[0..max) and
 * returns the count, so bindings can check their mirrors (host only). */
int32_t pdlp_abi_layout(int64_t *out, int32_t max);

#ifdef __cplusplus
}
#endif

#endif /* PDLP_HIP_H */
