"""Writes tests/golden/zipf_scaled_oracle.json: the oracle's FP64 solve of
the Zipf-degree LP generator at n = 20,000 (the n = 1,000,000 benchmark LP's
generator, scale 0.02) to 1e-4 -- the termination reason, iteration count
and final objectives -- so GPU tests can compare a long solve (about 474k
iterations, 5 minutes of the C oracle) without running it.

Run from the repository root:  python tests/golden/make_zipf_oracle.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "cupdlp.jl_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402
from cupdlp import PdhgParameters, construct_termination_criteria  # noqa: E402
from cupdlp.generators import config_lp  # noqa: E402

TOL = 1e-4


def main():
    prob, _ = config_lp(1, scale=0.02)
    tc = construct_termination_criteria(eps_optimal_absolute=TOL, eps_optimal_relative=TOL)
    t = time.time()
    out = O.solve(prob, PdhgParameters(termination_criteria=tc))
    ci = out.iteration_stats[-1].convergence_information[0]
    rec = {"lp": "config_lp(1, scale=0.02)", "tolerance": TOL,
           "termination_reason": int(out.termination_reason),
           "iteration_count": int(out.iteration_count),
           "primal_objective": ci.primal_objective, "dual_objective": ci.dual_objective,
           "oracle_seconds": round(time.time() - t, 1)}
    print(rec, flush=True)
    with open(os.path.join(HERE, "zipf_scaled_oracle.json"), "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
