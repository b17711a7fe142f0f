"""The command-line solve (cupdlp.solve, after scripts/solve.jl): instance
names, the SolveLog JSON (src/solve_log.jl:336-402), the extension check,
and on the GPU the four output files of a solve against a direct solve."""
import gzip
import json
import math
import os

import numpy as np
import pytest

import mps_oracle as MO
from cupdlp import (IterationStats, PdhgParameters, SolveLog, TerminationReason,
                    construct_termination_criteria, optimize, qps_reader_to_standard_form)
from cupdlp.generators import random_lp
from cupdlp.solve import instance_name, main, solve_instance_and_output, solve_log_json
from cupdlp.types import ConvergenceInformation, PointType, RestartChoice


def test_instance_name():
    # scripts/solve.jl:20: the .mps / .qps extension (optionally .gz) is dropped
    assert instance_name("/a/b/afiro.mps.gz") == "afiro"
    assert instance_name("afiro.MPS") == "afiro"
    assert instance_name("x.qps") == "x"
    assert instance_name("x.mps.bz2") == "x.mps.bz2"


def test_solve_log_json_fields():
    st = IterationStats(iteration_number=7, convergence_information=[ConvergenceInformation(
        candidate_type=PointType.POINT_TYPE_AVERAGE_ITERATE, primal_objective=1.5,
        l_inf_primal_variable=math.inf)], cumulative_time_sec=0.25,
        restart_used=RestartChoice.RESTART_CHOICE_NO_RESTART,
        method_specific_stats={"time_spent_doing_basic_algorithm": 0.2, "total_attempts": 9.0})
    log = SolveLog(instance_name="t", command_line_invocation="solve.py --x",
                   termination_reason=TerminationReason.TERMINATION_REASON_OPTIMAL,
                   termination_string="OPTIMAL", iteration_count=7, solve_time_sec=0.25,
                   solution_stats=st, solution_type=PointType.POINT_TYPE_AVERAGE_ITERATE)
    d = json.loads(solve_log_json(log))
    assert list(d) == ["instance_name", "command_line_invocation", "termination_reason",
                       "termination_string", "iteration_count", "solve_time_sec",
                       "solution_stats", "solution_type", "iteration_stats"]
    assert d["termination_reason"] == "TERMINATION_REASON_OPTIMAL"
    assert d["solution_type"] == "POINT_TYPE_AVERAGE_ITERATE"
    s = d["solution_stats"]
    assert list(s) == ["iteration_number", "convergence_information", "infeasibility_information",
                       "cumulative_kkt_matrix_passes", "cumulative_rejected_steps",
                       "cumulative_time_sec", "restart_used", "step_size", "primal_weight",
                       "method_specific_stats"]
    assert s["restart_used"] == "RESTART_CHOICE_NO_RESTART"
    assert s["method_specific_stats"] == {"time_spent_doing_basic_algorithm": 0.2}
    ci = s["convergence_information"][0]
    assert ci["candidate_type"] == "POINT_TYPE_AVERAGE_ITERATE"
    assert ci["primal_objective"] == 1.5 and ci["l_inf_primal_variable"] == math.inf
    assert d["iteration_stats"] == []


def test_unrecognized_extension(tmp_path):
    with pytest.raises(ValueError):
        solve_instance_and_output(PdhgParameters(), str(tmp_path), str(tmp_path / "x.lp"))


@pytest.mark.gpu
def test_cli_solve_writes_outputs(gpu_ctx, tmp_path):
    print("case: command-line solve of an n=2000 LP written as MPS (tolerance 1e-6)", flush=True)
    prob = random_lp(n=2000, m=1000, num_eq=500, nnz_per_col=8, seed=4)[0]
    path = os.path.join(tmp_path, "gen4.mps")
    MO.write_free_mps(path, prob.objective_vector, 0.0, prob.variable_lower_bound,
                      prob.variable_upper_bound, prob.constraint_matrix, prob.right_hand_side,
                      prob.num_equalities)
    out_dir = os.path.join(tmp_path, "out")
    assert main(["--instance_path", path, "--output_directory", out_dir, "--tolerance", "1e-6"]) == 0
    with open(os.path.join(out_dir, "gen4_summary.json")) as f:
        summary = json.load(f)
    assert summary["instance_name"] == "gen4"
    assert summary["termination_reason"] == "TERMINATION_REASON_OPTIMAL"
    assert summary["iteration_stats"] == []
    with gzip.open(os.path.join(out_dir, "gen4_full_log.json.gz"), "rt") as f:
        full = json.load(f)
    assert len(full["iteration_stats"]) > 0
    assert full["iteration_count"] == summary["iteration_count"]
    x = np.loadtxt(os.path.join(out_dir, "gen4_primal.txt"))
    y = np.loadtxt(os.path.join(out_dir, "gen4_dual.txt"))
    # the same solve through the API: the same point, bitwise
    tc = construct_termination_criteria(eps_optimal_absolute=1e-6, eps_optimal_relative=1e-6,
                                        time_sec_limit=3600.0)
    ref = optimize(PdhgParameters(termination_criteria=tc), qps_reader_to_standard_form(path))
    assert ref.iteration_count == summary["iteration_count"]
    assert np.array_equal(x, ref.primal_solution) and np.array_equal(y, ref.dual_solution)
