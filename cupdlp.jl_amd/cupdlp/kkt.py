"""The reference's optimality measures of a returned point, on the host.

``relative_kkt(problem, x, y)`` is the largest of the three quantities the
reference's optimality test compares with eps when eps_optimal_absolute =
eps_optimal_relative = eps (src/termination.jl:142-170, L2 norm): the
primal residual over 1 + ||b||_2, the dual residual over 1 + ||c||_2 and the
duality gap over 1 + |primal objective| + |dual objective|, each built as in
src/iteration_stats_utils_gpu.jl:26-309 (">=" rows clipped, bound
violations, reduced costs projected by finite bounds, the dual objective's
reduced-cost term).  A point meets the test at eps iff the value is < eps.
Computed with numpy in plain FP64, for reporting (benchmarks, tests); the
solver's own statistics come from the device.
"""
import numpy as np


def kkt_parts(problem, x, y):
    """(relative primal residual, relative dual residual, relative gap,
    primal objective, dual objective) of (x, y)."""
    A = problem.constraint_matrix
    neq = problem.num_equalities
    b, c = problem.right_hand_side, problem.objective_vector
    lo, hi = problem.variable_lower_bound, problem.variable_upper_bound
    t = b - A @ x
    pv = np.concatenate([t[:neq], np.maximum(t[neq:], 0.0), np.maximum(lo - x, 0.0),
                         np.maximum(x - hi, 0.0)])
    g = c - A.T @ y
    flo, fhi = np.isfinite(lo), np.isfinite(hi)
    lam = np.maximum(g, 0.0) * flo + np.minimum(g, 0.0) * fhi
    dv = np.concatenate([np.maximum(-y[neq:], 0.0), g - lam])
    p = problem.objective_constant + c @ x
    contrib = np.where(lam > 0, np.where(flo, lo, 0.0) * lam,
                       np.where(lam < 0, np.where(fhi, hi, 0.0) * lam, 0.0))
    d = problem.objective_constant + b @ y + contrib.sum()
    return (np.linalg.norm(pv) / (1.0 + np.linalg.norm(b)),
            np.linalg.norm(dv) / (1.0 + np.linalg.norm(c)),
            abs(p - d) / (1.0 + abs(p) + abs(d)), p, d)


def relative_kkt(problem, x, y):
    rp, rd, rg, _, _ = kkt_parts(problem, x, y)
    return max(rp, rd, rg)
