"""Host-side mirror of cuPDLP.jl's public types.

Names, field names, field order, defaults and enum values follow the
reference (paths relative to the cuPDLP.jl tree):

* ``QuadraticProgrammingProblem``   src/quadratic_programming.jl:25-67
* ``PdhgParameters``                src/primal_dual_hybrid_gradient_gpu.jl:9-24
* ``AdaptiveStepsizeParams`` / ``ConstantStepsizeParams``  ...gpu.jl:2-7
* ``TerminationCriteria`` + ``construct_termination_criteria``  src/termination.jl:9-99
* ``RestartParameters`` + ``construct_restart_parameters``  src/saddle_point_gpu.jl:353-414
* ``ConvergenceInformation`` / ``InfeasibilityInformation`` / ``IterationStats``
  src/solve_log.jl:51-302
* ``SaddlePointOutput``              src/saddle_point_gpu.jl:3-34
* ``SolveLog``                       src/solve_log.jl:336-402
* the enums                         src/solve_log.jl:19-45, 323-334;
  src/saddle_point_gpu.jl:343-350; src/termination.jl:6
"""
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional, Union

import numpy as np
import scipy.sparse as sp


class TerminationReason(IntEnum):
    TERMINATION_REASON_UNSPECIFIED = 0
    TERMINATION_REASON_OPTIMAL = 1
    TERMINATION_REASON_PRIMAL_INFEASIBLE = 2
    TERMINATION_REASON_DUAL_INFEASIBLE = 3
    TERMINATION_REASON_TIME_LIMIT = 4
    TERMINATION_REASON_ITERATION_LIMIT = 5
    TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT = 6
    TERMINATION_REASON_NUMERICAL_ERROR = 7
    TERMINATION_REASON_INVALID_PROBLEM = 8
    TERMINATION_REASON_OTHER = 9


class RestartChoice(IntEnum):
    RESTART_CHOICE_UNSPECIFIED = 0
    RESTART_CHOICE_NO_RESTART = 1
    RESTART_CHOICE_WEIGHTED_AVERAGE_RESET = 2
    RESTART_CHOICE_RESTART_TO_AVERAGE = 3


class PointType(IntEnum):
    POINT_TYPE_UNSPECIFIED = 0
    POINT_TYPE_CURRENT_ITERATE = 1
    POINT_TYPE_ITERATE_DIFFERENCE = 2
    POINT_TYPE_AVERAGE_ITERATE = 3
    POINT_TYPE_NONE = 4


class RestartScheme(IntEnum):
    NO_RESTARTS = 0
    FIXED_FREQUENCY = 1
    ADAPTIVE_KKT = 2


class RestartToCurrentMetric(IntEnum):
    NO_RESTART_TO_CURRENT = 0
    KKT_GREEDY = 1


class OptimalityNorm(IntEnum):
    L_INF = 0
    L2 = 1


class Precision(IntEnum):
    """Arithmetic of the bulk iteration (extension)."""
    FP64 = 0
    FP32_REFINE = 1


# Module-level aliases, as Julia exports @enum values into the module.
NO_RESTARTS = RestartScheme.NO_RESTARTS
FIXED_FREQUENCY = RestartScheme.FIXED_FREQUENCY
ADAPTIVE_KKT = RestartScheme.ADAPTIVE_KKT
NO_RESTART_TO_CURRENT = RestartToCurrentMetric.NO_RESTART_TO_CURRENT
KKT_GREEDY = RestartToCurrentMetric.KKT_GREEDY
L_INF = OptimalityNorm.L_INF
L2 = OptimalityNorm.L2


def termination_reason_to_string(reason) -> str:
    """Strips the TERMINATION_REASON_ prefix (src/termination.jl:248-251)."""
    return TerminationReason(int(reason)).name[len("TERMINATION_REASON_"):]


@dataclass
class QuadraticProgrammingProblem:
    """min c'x + c0  s.t. A[:num_eq] x = b[:num_eq], A[num_eq:] x >= b[num_eq:],
    l <= x <= u.  The objective matrix must be zero (the reference's GPU path
    ignores it, src/cpu_to_gpu.jl:1-14)."""
    variable_lower_bound: np.ndarray
    variable_upper_bound: np.ndarray
    objective_matrix: Optional[sp.spmatrix]
    objective_vector: np.ndarray
    objective_constant: float
    constraint_matrix: sp.spmatrix
    right_hand_side: np.ndarray
    num_equalities: int

    def __post_init__(self):
        self.variable_lower_bound = np.ascontiguousarray(self.variable_lower_bound, dtype=np.float64)
        self.variable_upper_bound = np.ascontiguousarray(self.variable_upper_bound, dtype=np.float64)
        self.objective_vector = np.ascontiguousarray(self.objective_vector, dtype=np.float64)
        self.right_hand_side = np.ascontiguousarray(self.right_hand_side, dtype=np.float64)
        A = sp.csc_matrix(self.constraint_matrix, dtype=np.float64)
        A.sum_duplicates()
        A.sort_indices()
        self.constraint_matrix = A
        self.objective_constant = float(self.objective_constant)
        self.num_equalities = int(self.num_equalities)
        if self.objective_matrix is not None and sp.csc_matrix(self.objective_matrix).nnz != 0:
            raise ValueError("quadratic objectives are not supported by the GPU path")

    @property
    def num_variables(self) -> int:
        return self.constraint_matrix.shape[1]

    @property
    def num_constraints(self) -> int:
        return self.constraint_matrix.shape[0]


def linear_programming_problem(variable_lower_bound, variable_upper_bound, objective_vector,
                               objective_constant, constraint_matrix, right_hand_side,
                               num_equalities) -> QuadraticProgrammingProblem:
    """src/quadratic_programming.jl:246-268."""
    return QuadraticProgrammingProblem(variable_lower_bound, variable_upper_bound, None,
                                       objective_vector, objective_constant, constraint_matrix,
                                       right_hand_side, num_equalities)


@dataclass
class AdaptiveStepsizeParams:
    reduction_exponent: float = 0.3
    growth_exponent: float = 0.6


@dataclass
class ConstantStepsizeParams:
    pass


@dataclass
class TerminationCriteria:
    optimality_norm: OptimalityNorm = OptimalityNorm.L2
    eps_optimal_absolute: float = 1.0e-6
    eps_optimal_relative: float = 1.0e-6
    eps_primal_infeasible: float = 1.0e-8
    eps_dual_infeasible: float = 1.0e-8
    time_sec_limit: float = float("inf")
    iteration_limit: int = 2**31 - 1
    kkt_matrix_pass_limit: float = float("inf")


def construct_termination_criteria(*, optimality_norm=OptimalityNorm.L2,
                                   eps_optimal_absolute=1.0e-6, eps_optimal_relative=1.0e-6,
                                   eps_primal_infeasible=1.0e-8, eps_dual_infeasible=1.0e-8,
                                   time_sec_limit=float("inf"), iteration_limit=2**31 - 1,
                                   kkt_matrix_pass_limit=float("inf")) -> TerminationCriteria:
    """src/termination.jl:79-99."""
    return TerminationCriteria(optimality_norm, eps_optimal_absolute, eps_optimal_relative,
                               eps_primal_infeasible, eps_dual_infeasible, time_sec_limit,
                               iteration_limit, kkt_matrix_pass_limit)


@dataclass
class RestartParameters:
    restart_scheme: RestartScheme = RestartScheme.ADAPTIVE_KKT
    restart_to_current_metric: RestartToCurrentMetric = RestartToCurrentMetric.KKT_GREEDY
    restart_frequency_if_fixed: int = 1000
    artificial_restart_threshold: float = 0.36
    sufficient_reduction_for_restart: float = 0.2
    necessary_reduction_for_restart: float = 0.8
    primal_weight_update_smoothing: float = 0.5


def construct_restart_parameters(restart_scheme, restart_to_current_metric,
                                 restart_frequency_if_fixed, artificial_restart_threshold,
                                 sufficient_reduction_for_restart,
                                 necessary_reduction_for_restart,
                                 primal_weight_update_smoothing) -> RestartParameters:
    """src/saddle_point_gpu.jl:388-414 (same assertions, raised as AssertionError)."""
    assert restart_frequency_if_fixed > 1
    assert 0.0 < artificial_restart_threshold <= 1.0
    assert 0.0 < sufficient_reduction_for_restart <= necessary_reduction_for_restart <= 1.0
    assert 0.0 <= primal_weight_update_smoothing <= 1.0
    return RestartParameters(RestartScheme(restart_scheme),
                             RestartToCurrentMetric(restart_to_current_metric),
                             int(restart_frequency_if_fixed), artificial_restart_threshold,
                             sufficient_reduction_for_restart, necessary_reduction_for_restart,
                             primal_weight_update_smoothing)


@dataclass
class PdhgParameters:
    """Positional field order of src/primal_dual_hybrid_gradient_gpu.jl:9-24;
    the keyword-only tail holds this engine's extensions."""
    l_inf_ruiz_iterations: int = 10
    l2_norm_rescaling: bool = False
    pock_chambolle_alpha: Optional[float] = 1.0
    primal_importance: float = 1.0
    scale_invariant_initial_primal_weight: bool = True
    verbosity: int = 2
    record_iteration_stats: bool = True
    termination_evaluation_frequency: int = 64
    termination_criteria: TerminationCriteria = field(default_factory=TerminationCriteria)
    restart_params: RestartParameters = field(default_factory=RestartParameters)
    step_size_policy_params: Union[AdaptiveStepsizeParams, ConstantStepsizeParams] = field(
        default_factory=AdaptiveStepsizeParams)
    # --- extensions (no reference counterpart) ---
    precision: Precision = Precision.FP64
    refine_max_outer: int = 12
    refine_inner_tolerance: float = 1.0e-4
    refine_max_scale: float = 1.0e6
    device_preconditioning: bool = True
    refine_inner_iteration_limit: int = 0
    batch_groups: int = 0
    refine_early_fallback: int = 0
    debug_numerical_error_attempt: int = 0  # test-only fault injection (movement = 0)
    refine_warm_start: int = 3
    lane_budget: int = 0    # step-kernel tiles (0 = default; see pdlp_tune)
    slice_unroll: int = 0
    slice_waves: int = 0
    primal_occupancy: int = 0  # most waves per SIMD of the slice kernels (0: the
    dual_occupancy: int = 0    # defaults 5 / 3; 8 or more: no cap)
    refine_continue_progress: int = 0  # > 0: a capped correction solve that cut the
                                       # error this many times continues the refinement


@dataclass
class ConvergenceInformation:
    candidate_type: PointType = PointType.POINT_TYPE_UNSPECIFIED
    primal_objective: float = 0.0
    dual_objective: float = 0.0
    corrected_dual_objective: float = 0.0
    l_inf_primal_residual: float = 0.0
    l2_primal_residual: float = 0.0
    l_inf_dual_residual: float = 0.0
    l2_dual_residual: float = 0.0
    relative_l_inf_primal_residual: float = 0.0
    relative_l2_primal_residual: float = 0.0
    relative_l_inf_dual_residual: float = 0.0
    relative_l2_dual_residual: float = 0.0
    relative_optimality_gap: float = 0.0
    l_inf_primal_variable: float = 0.0
    l2_primal_variable: float = 0.0
    l_inf_dual_variable: float = 0.0
    l2_dual_variable: float = 0.0


@dataclass
class InfeasibilityInformation:
    candidate_type: PointType = PointType.POINT_TYPE_UNSPECIFIED
    max_primal_ray_infeasibility: float = 0.0
    primal_ray_linear_objective: float = 0.0
    max_dual_ray_infeasibility: float = 0.0
    dual_ray_objective: float = 0.0


@dataclass
class IterationStats:
    iteration_number: int = 0
    convergence_information: List[ConvergenceInformation] = field(default_factory=list)
    infeasibility_information: List[InfeasibilityInformation] = field(default_factory=list)
    cumulative_kkt_matrix_passes: float = 0.0
    cumulative_rejected_steps: int = 0
    cumulative_time_sec: float = 0.0
    restart_used: RestartChoice = RestartChoice.RESTART_CHOICE_UNSPECIFIED
    step_size: float = 0.0
    primal_weight: float = 0.0
    method_specific_stats: dict = field(default_factory=dict)


@dataclass
class SolveLog:
    instance_name: str = ""
    command_line_invocation: str = ""
    termination_reason: TerminationReason = TerminationReason.TERMINATION_REASON_UNSPECIFIED
    termination_string: str = ""
    iteration_count: int = 0
    solve_time_sec: float = 0.0
    solution_stats: IterationStats = field(default_factory=IterationStats)
    solution_type: PointType = PointType.POINT_TYPE_UNSPECIFIED
    iteration_stats: List[IterationStats] = field(default_factory=list)


@dataclass
class SaddlePointOutput:
    primal_solution: np.ndarray
    dual_solution: np.ndarray
    termination_reason: TerminationReason
    termination_string: str
    iteration_count: int
    iteration_stats: List[IterationStats]
    # extensions
    solve_time_sec: float = 0.0
    preprocess_time_sec: float = 0.0
    total_attempts: int = 0
    refinement_rounds: int = 0
    fp64_fallback: bool = False
    primal_current: Optional[np.ndarray] = None
    dual_current: Optional[np.ndarray] = None
