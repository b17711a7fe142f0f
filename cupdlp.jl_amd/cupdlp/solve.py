"""Command-line solve of an MPS / QPS file, as scripts/solve.jl does.

    python -m cupdlp.solve --instance_path X.mps[.gz] --output_directory D
                           [--tolerance 1e-4] [--time_sec_limit 3600]

Writes, like solve_instance_and_output (scripts/solve.jl:14-78):
<name>_summary.json (the SolveLog of src/solve_log.jl:336-402 without the
iteration history), <name>_full_log.json.gz (with it), <name>_primal.txt and
<name>_dual.txt (one value per line).  Enumerations are written by name;
floats in shortest round-trip form (a reader gets the same Float64 values);
non-finite values as Infinity / -Infinity / NaN.
"""
import argparse
import dataclasses
import enum
import gzip
import json
import math
import os
import re
import sys

from .io import qps_reader_to_standard_form
from .solver import optimize
from .types import (AdaptiveStepsizeParams, PdhgParameters, PointType, SolveLog,
                    construct_restart_parameters, construct_termination_criteria)
from .types import ADAPTIVE_KKT, KKT_GREEDY

_EXT = re.compile(r"\.(mps|MPS|qps|QPS)(\.gz)?$")


def _jsonable(v):
    if isinstance(v, enum.Enum):
        return v.name
    if dataclasses.is_dataclass(v):
        return {f.name: _jsonable(getattr(v, f.name)) for f in dataclasses.fields(v)}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, dict):
        # the reference's method-specific statistics hold this one entry
        return {k: _jsonable(x) for k, x in v.items() if k == "time_spent_doing_basic_algorithm"}
    if isinstance(v, float):
        return float(v)
    if hasattr(v, "item"):  # numpy scalars
        return v.item()
    return v


def solve_log_json(log):
    return json.dumps(_jsonable(log), allow_nan=True)


def _write_vector(path, vec):
    with open(path, "w") as f:
        for x in vec:
            x = float(x)
            f.write(("Inf" if x > 0 else "-Inf") if math.isinf(x) else ("NaN" if math.isnan(x) else repr(x)))
            f.write("\n")


def instance_name(instance_path):
    return _EXT.sub("", os.path.basename(instance_path))


def solve_instance_and_output(parameters, output_dir, instance_path, argv=None):
    """scripts/solve.jl:14-78."""
    os.makedirs(output_dir, exist_ok=True)
    name = instance_name(instance_path)
    low = os.path.basename(instance_path).lower()
    if not low.endswith((".mps", ".mps.gz", ".qps", ".qps.gz")):
        raise ValueError(f"Instance has unrecognized file extension: {os.path.basename(instance_path)}")
    lp = qps_reader_to_standard_form(instance_path)
    if parameters.verbosity >= 1:
        print("Instance:", name)
    output = optimize(parameters, lp)
    last = output.iteration_stats[-1]
    log = SolveLog(instance_name=name,
                   command_line_invocation=" ".join(argv if argv is not None else sys.argv),
                   termination_reason=output.termination_reason,
                   termination_string=output.termination_string,
                   iteration_count=output.iteration_count,
                   solve_time_sec=last.cumulative_time_sec, solution_stats=last,
                   solution_type=PointType.POINT_TYPE_AVERAGE_ITERATE)
    with open(os.path.join(output_dir, name + "_summary.json"), "w") as f:
        f.write(solve_log_json(log))
    log.iteration_stats = output.iteration_stats
    with gzip.open(os.path.join(output_dir, name + "_full_log.json.gz"), "wt") as f:
        f.write(solve_log_json(log))
    _write_vector(os.path.join(output_dir, name + "_primal.txt"), output.primal_solution)
    _write_vector(os.path.join(output_dir, name + "_dual.txt"), output.dual_solution)
    return output


def _restart_params():
    return construct_restart_parameters(ADAPTIVE_KKT, KKT_GREEDY, 1000, 0.36, 0.2, 0.8, 0.5)


def warm_up(lp):
    """scripts/solve.jl:80-119: a 100-iteration solve before the timed one."""
    tc = construct_termination_criteria(eps_optimal_absolute=1.0e-4, eps_optimal_relative=1.0e-4,
                                        eps_primal_infeasible=1.0e-8, eps_dual_infeasible=1.0e-8,
                                        time_sec_limit=math.inf, iteration_limit=100,
                                        kkt_matrix_pass_limit=math.inf)
    optimize(PdhgParameters(10, False, 1.0, 1.0, True, 0, True, 64, tc, _restart_params(),
                            AdaptiveStepsizeParams(0.3, 0.6)), lp)


def main(argv=None):
    """scripts/solve.jl:121-206."""
    argv = list(sys.argv[1:] if argv is None else argv)
    ap = argparse.ArgumentParser(prog="python -m cupdlp.solve")
    ap.add_argument("--instance_path", required=True,
                    help="The path to the instance to solve in .mps.gz or .mps format.")
    ap.add_argument("--output_directory", required=True, help="The directory for output files.")
    ap.add_argument("--tolerance", type=float, default=1e-4, help="KKT tolerance of the solution.")
    ap.add_argument("--time_sec_limit", type=float, default=3600.0, help="Time limit.")
    a = ap.parse_args(argv)
    lp = qps_reader_to_standard_form(a.instance_path)
    warm_up(lp)
    tc = construct_termination_criteria(eps_optimal_absolute=a.tolerance,
                                        eps_optimal_relative=a.tolerance,
                                        eps_primal_infeasible=1.0e-8, eps_dual_infeasible=1.0e-8,
                                        time_sec_limit=a.time_sec_limit, iteration_limit=2**31 - 1,
                                        kkt_matrix_pass_limit=math.inf)
    params = PdhgParameters(10, False, 1.0, 1.0, True, 2, True, 64, tc, _restart_params(),
                            AdaptiveStepsizeParams(0.3, 0.6))
    solve_instance_and_output(params, a.output_directory, a.instance_path,
                              argv=["solve.py"] + argv)
    return 0


if __name__ == "__main__":
    sys.exit(main())
