"""cupdlp — host-side mirror of cuPDLP.jl over the MI355X C ABI (libpdlp_hip.so).

The public surface follows the reference module (src/cuPDLP.jl): ``optimize``,
``PdhgParameters``, ``QuadraticProgrammingProblem``, the restart / termination
constructors, the enums, ``SaddlePointOutput`` and ``Optimizer`` (the model
front end of src/MOI_wrapper.jl).
"""
from .types import *  # noqa: F401,F403
from .types import (AdaptiveStepsizeParams, ConstantStepsizeParams, PdhgParameters,
                    QuadraticProgrammingProblem, SaddlePointOutput, TerminationCriteria,
                    RestartParameters, construct_restart_parameters,
                    construct_termination_criteria, linear_programming_problem,
                    termination_reason_to_string)
from .solver import (Context, LoadedBatch, PdlpError, Server, default_context,  # noqa: F401
                     default_params, optimize)
from .moi import Optimizer  # noqa: F401
from .io import MpsError, qps_reader_to_standard_form  # noqa: F401
from .types import SolveLog  # noqa: F401
from .presolve import PresolveInfo, presolve, undo_presolve  # noqa: F401
