"""Writes tests/golden/highs_objectives.json: optimal objectives of the test
LPs computed by HiGHS (scipy.optimize.linprog, method="highs") in FP64, an
independent simplex/IPM solver used to pin the oracle's end-to-end answers.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np
import scipy
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "cupdlp.jl_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import lp_cases  # noqa: E402
from cupdlp.generators import config_lp, random_lp, transportation_lp  # noqa: E402


def highs(prob):
    A = prob.constraint_matrix.tocsr()
    k = prob.num_equalities
    bounds = [(None if not np.isfinite(l) else l, None if not np.isfinite(u) else u)
              for l, u in zip(prob.variable_lower_bound, prob.variable_upper_bound)]
    kw = {}
    if k > 0:
        kw["A_eq"] = A[:k]
        kw["b_eq"] = prob.right_hand_side[:k]
    if A.shape[0] > k:
        kw["A_ub"] = -A[k:]
        kw["b_ub"] = -prob.right_hand_side[k:]
    r = linprog(prob.objective_vector, bounds=bounds, method="highs", **kw)
    assert r.status == 0, r.message
    return float(r.fun) + prob.objective_constant


CASES = {
    "c0_seed0": lambda: config_lp(0)[0],
    "random_n300_s11": lambda: random_lp(n=300, m=150, seed=11)[0],
    "random_n1200_s12": lambda: random_lp(n=1200, m=600, seed=12)[0],
    "random_n400_s123": lambda: random_lp(n=400, m=200, num_eq=100, seed=123)[0],
    "transport_60x60_s5": lambda: transportation_lp(60, 60, seed=5),
    "transport_200x200_s2": lambda: transportation_lp(200, 200, seed=2),
    "long_row_6000": lambda: lp_cases.long_row_lp(6000)[0],
    "zipf_scaled_n20000": lambda: config_lp(1, scale=0.02)[0],
    # the 2,000 x 2,000 transportation LP (n = 4,000,000); about 2 minutes
    "transport_2000x2000_s2": lambda: config_lp(2)[0],
}


def main():
    out = {"solver": f"HiGHS via scipy {scipy.__version__}", "objectives": {}}
    for name, mk in CASES.items():
        prob = mk()
        out["objectives"][name] = {"objective": highs(prob), "n": prob.num_variables,
                                   "m": prob.num_constraints,
                                   "nnz": int(prob.constraint_matrix.nnz)}
        print(name, out["objectives"][name], flush=True)
    with open(os.path.join(HERE, "highs_objectives.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
