"""Seeded synthetic LP generators with certified optima.

``random_lp`` builds an LP together with a primal-dual optimal pair
(x*, y*, reduced costs r) by construction:

* bounded variables (0 <= x <= 10) and free variables; 40 % of the bounded
  ones sit at a bound (half at the lower, half at the upper bound), the rest
  are U(0.5, 9.5); free ones are N(0, 1);
* equality rows carry y* ~ N(0, 1); ">=" rows are half slack (y* = 0,
  slack ~ U(0.1, 1)) and half tight (y* ~ U(0.1, 1), slack 0);
* b = A x* - slack and c = A^T y* + r, with r ~ U(0.1, 1) positive at a
  lower bound, negative at an upper bound and zero elsewhere.

Then x* is primal feasible, y* dual feasible, and complementary slackness
holds, so c'x* is the optimal value.  Optional row/column scaling
A' = R A C keeps the certificate (x' = x*/C, y' = y*/R, b' = R b, c' = C c).

``transportation_lp`` builds the supply/demand LP with exactly balanced
integer supplies and demands.

All randomness comes from numpy's PCG64 with the given seed.
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .types import QuadraticProgrammingProblem


@dataclass
class Certificate:
    x: np.ndarray
    y: np.ndarray
    r: np.ndarray
    objective: float


def _distinct_rows(rng, cols, m, weights=None, max_rounds=20):
    """Row indices for the entries of each column (cols sorted), redrawing
    duplicates within a column; leftover duplicates are dropped."""
    nnz = len(cols)
    if weights is None:
        rows = rng.integers(0, m, size=nnz)
    else:
        cdf = np.cumsum(weights); cdf /= cdf[-1]
        rows = np.minimum(np.searchsorted(cdf, rng.random(nnz), side="right"), m - 1)
    for _ in range(max_rounds):
        key = cols.astype(np.int64) * m + rows
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(nnz, dtype=bool)
        dup[order[1:][ks[1:] == ks[:-1]]] = True
        k = int(dup.sum())
        if k == 0:
            break
        if weights is None:
            rows[dup] = rng.integers(0, m, size=k)
        else:
            rows[dup] = np.minimum(np.searchsorted(cdf, rng.random(k), side="right"), m - 1)
    key = cols.astype(np.int64) * m + rows
    _, keep = np.unique(key, return_index=True)
    keep.sort()
    return cols[keep], rows[keep]


def random_lp(n=2000, m=1000, num_eq=None, nnz_per_col=8, frac_free=0.1, seed=0,
              col_degrees=None, row_popularity_exponent=None, row_scale_decades=0.0,
              col_scale_decades=0.0):
    """Returns (QuadraticProgrammingProblem, Certificate)."""
    rng = np.random.default_rng(seed)
    if num_eq is None:
        num_eq = m // 2
    if col_degrees is None:
        deg = np.full(n, nnz_per_col, dtype=np.int64)
    else:
        deg = np.asarray(col_degrees, dtype=np.int64)
    deg = np.minimum(deg, m)
    cols = np.repeat(np.arange(n, dtype=np.int64), deg)
    weights = None
    if row_popularity_exponent is not None:
        ranks = rng.permutation(m).astype(np.float64) + 1.0
        weights = ranks ** (-row_popularity_exponent)
    cols, rows = _distinct_rows(rng, cols, m, weights)
    vals = rng.standard_normal(len(cols))
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    A.sort_indices()

    # variables: bounded [0, 10] or free
    n_free = int(round(frac_free * n))
    perm = rng.permutation(n)
    free = np.zeros(n, dtype=bool); free[perm[:n_free]] = True
    lb = np.where(free, -np.inf, 0.0)
    ub = np.where(free, np.inf, 10.0)
    bounded = np.flatnonzero(~free)
    nb = len(bounded)
    at_bound = rng.random(nb) < 0.4
    at_lower = rng.random(nb) < 0.5
    x = np.zeros(n)
    r = np.zeros(n)
    xb = rng.uniform(0.5, 9.5, size=nb)
    rb = rng.uniform(0.1, 1.0, size=nb)
    xb = np.where(at_bound, np.where(at_lower, 0.0, 10.0), xb)
    rb = np.where(at_bound, np.where(at_lower, rb, -rb), 0.0)
    x[bounded] = xb
    r[bounded] = rb
    x[free] = rng.standard_normal(n_free)

    # rows: equalities then ">=" rows
    y = np.zeros(m)
    slack = np.zeros(m)
    y[:num_eq] = rng.standard_normal(num_eq)
    n_ineq = m - num_eq
    tight = rng.random(n_ineq) < 0.5
    y_in = np.where(tight, rng.uniform(0.1, 1.0, size=n_ineq), 0.0)
    s_in = np.where(tight, 0.0, rng.uniform(0.1, 1.0, size=n_ineq))
    y[num_eq:] = y_in
    slack[num_eq:] = s_in

    b = A @ x - slack
    c = A.T @ y + r

    if row_scale_decades > 0 or col_scale_decades > 0:
        R = 10.0 ** rng.uniform(-row_scale_decades, row_scale_decades, size=m)
        Cs = 10.0 ** rng.uniform(-col_scale_decades, col_scale_decades, size=n)
        A = sp.csc_matrix(sp.diags(R) @ A @ sp.diags(Cs))
        A.sort_indices()
        b = R * b
        c = Cs * c
        lb = lb / Cs
        ub = ub / Cs
        x = x / Cs
        y = y / R
        r = Cs * r

    prob = QuadraticProgrammingProblem(lb, ub, None, c, 0.0, A, b, num_eq)
    cert = Certificate(x=x, y=y, r=r, objective=float(c @ x))
    return prob, cert


def zipf_degrees(rng, n, a=1.5, cap=2000, target_nnz=None):
    """Column degrees from Zipf(a) capped at `cap`, optionally rescaled to a
    total of `target_nnz` with stochastic rounding and a minimum of 1."""
    d = np.minimum(rng.zipf(a, size=n), cap).astype(np.float64)
    if target_nnz is not None:
        d = d * (target_nnz / d.sum())
        fl = np.floor(d)
        d = fl + (rng.random(n) < (d - fl))
        d = np.maximum(d, 1.0)
    return d.astype(np.int64)


def config_lp(index, seed=None, scale=1.0):
    """The benchmark configurations.  index 0: n=2,000 / m=1,000 random LP;
    index 1: n=1,000,000 / m=500,000 Zipf-degree scaled LP (~12M nnz);
    index 2: 2,000 x 2,000 transportation LP.  ``scale`` shrinks sizes for
    quick tests (1.0 = full size)."""
    if index == 0:
        s = 0 if seed is None else seed
        n = max(int(2000 * scale), 4)
        return random_lp(n=n, m=n // 2, num_eq=n // 4, seed=s)
    if index == 1:
        s = 1 if seed is None else seed
        n = max(int(1_000_000 * scale), 8)
        m = n // 2
        rng = np.random.default_rng(s + 1_000_003)
        deg = zipf_degrees(rng, n, 1.5, min(2000, m), target_nnz=12 * n)
        return random_lp(n=n, m=m, num_eq=m // 2, seed=s, col_degrees=deg,
                         row_popularity_exponent=0.5, row_scale_decades=3.0,
                         col_scale_decades=2.0)
    if index == 2:
        s = 2 if seed is None else seed
        k = max(int(2000 * scale), 2)
        return transportation_lp(k, k, seed=s), None
    raise ValueError(index)


def transportation_lp(n_src=2000, n_dst=2000, seed=2):
    """min sum c_ij x_ij s.t. sum_j x_ij = s_i, sum_i x_ij = d_j, x >= 0;
    integer supplies U{100..1000}, integer demands balanced exactly,
    costs U(1, 100); A is all +1 with two entries per column."""
    rng = np.random.default_rng(seed)
    supply = rng.integers(100, 1001, size=n_src).astype(np.float64)
    total = supply.sum()
    raw = rng.integers(100, 1001, size=n_dst).astype(np.float64)
    dem = np.floor(raw * (total / raw.sum()))
    rem = int(total - dem.sum())
    if rem > 0:
        np.add.at(dem, rng.choice(n_dst, size=rem, replace=True), 1.0)
    assert dem.sum() == total
    cost = rng.uniform(1.0, 100.0, size=n_src * n_dst)
    n = n_src * n_dst
    j = np.arange(n, dtype=np.int64)
    src = j // n_dst
    dst = j % n_dst
    rows = np.stack([src, n_src + dst], axis=1).reshape(-1)
    cols = np.repeat(j, 2)
    A = sp.csc_matrix((np.ones(2 * n), (rows, cols)), shape=(n_src + n_dst, n))
    A.sort_indices()
    b = np.concatenate([supply, dem])
    return QuadraticProgrammingProblem(np.zeros(n), np.full(n, np.inf), None, cost, 0.0, A, b,
                                       n_src + n_dst)


def batch_sizes(count=512, n_min=500, n_max=200_000, seed=3):
    """Log-uniform problem sizes of the batched workload (n; m = n // 2)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=count)).astype(np.int64)
