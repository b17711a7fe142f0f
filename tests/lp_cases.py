"""Small LPs with hand-derived answers, shared by the oracle and GPU tests."""
import numpy as np
import scipy.sparse as sp

from cupdlp import QuadraticProgrammingProblem

INF = np.inf


def lp(A, b, c, lb, ub, neq, c0=0.0):
    return QuadraticProgrammingProblem(np.asarray(lb, float), np.asarray(ub, float), None,
                                       np.asarray(c, float), c0, sp.csc_matrix(np.asarray(A, float)),
                                       np.asarray(b, float), neq)


def tiny_optimal():
    """min -x  s.t. -x >= -1, 0 <= x: optimum -1 at x = 1, y = 1."""
    return lp([[-1.0]], [-1.0], [-1.0], [0.0], [INF], 0), -1.0


def tiny_two_var():
    """min x1 + 2 x2 s.t. x1 + x2 = 2, x1 - x2 >= -1, x >= 0  -> x = (2, 0), obj 2."""
    return lp([[1, 1], [1, -1]], [2, -1], [1, 2], [0, 0], [INF, INF], 1), 2.0


def primal_infeasible():
    """x >= 1 (row) with x <= 0 (bound)."""
    return lp([[1.0]], [1.0], [1.0], [-INF], [0.0], 0)


def dual_infeasible():
    """min -x s.t. x >= 0 (row), x >= 0: unbounded."""
    return lp([[1.0]], [0.0], [-1.0], [0.0], [INF], 0)


def no_rows():
    """m = 0: min x1 - x2 with 0 <= x1 <= 1, -1 <= x2 <= 2 -> x = (0, 2), obj -2."""
    A = sp.csc_matrix((0, 2))
    return QuadraticProgrammingProblem(np.array([0.0, -1.0]), np.array([1.0, 2.0]), None,
                                       np.array([1.0, -1.0]), 0.0, A, np.zeros(0), 0), -2.0


def empty_row_and_column_lp(n=300, m=150, seed=31):
    """A random LP plus an empty equality row (b = 0, first), an empty ">="
    row (b = -1, last) and an empty column (cost 1 on [0, 5], so x = 0):
    the generator's certificate, extended by zeros, stays optimal."""
    from cupdlp.generators import random_lp
    prob, cert = random_lp(n=n, m=m, num_eq=m // 3, seed=seed)
    A = sp.csc_matrix(prob.constraint_matrix)
    A = sp.hstack([A, sp.csc_matrix((m, 1))]).tocsc()
    A = sp.vstack([sp.csr_matrix((1, n + 1)), A, sp.csr_matrix((1, n + 1))]).tocsc()
    b = np.concatenate([[0.0], prob.right_hand_side, [-1.0]])
    c = np.append(prob.objective_vector, 1.0)
    lb = np.append(prob.variable_lower_bound, 0.0)
    ub = np.append(prob.variable_upper_bound, 5.0)
    p = QuadraticProgrammingProblem(lb, ub, None, c, 0.0, A, b, prob.num_equalities + 1)
    return p, cert.objective


def fixed_variables_lp(n=300, m=150, seed=43, every=10):
    """A random LP with every `every`-th variable fixed (lb = ub) at its
    certified optimal value: the certificate stays optimal."""
    from cupdlp.generators import random_lp
    prob, cert = random_lp(n=n, m=m, seed=seed)
    lb, ub = prob.variable_lower_bound.copy(), prob.variable_upper_bound.copy()
    lb[::every] = cert.x[::every]
    ub[::every] = cert.x[::every]
    p = QuadraticProgrammingProblem(lb, ub, None, prob.objective_vector, 0.0,
                                    prob.constraint_matrix, prob.right_hand_side,
                                    prob.num_equalities)
    return p, cert


def one_sided_bounds_lp(n=300, m=150, seed=47):
    """A random LP whose [0, 10] variables lose the bound their certificate
    does not use: at 0 with r > 0 -> [0, inf); at 10 with r < 0 -> (-inf, 10];
    interior (r = 0) -> alternately (-inf, 10] and [0, inf).  Every bound the
    certificate's complementarity needs is kept, so it stays optimal."""
    from cupdlp.generators import random_lp
    prob, cert = random_lp(n=n, m=m, seed=seed)
    lb, ub = prob.variable_lower_bound.copy(), prob.variable_upper_bound.copy()
    x, r = cert.x, cert.r
    bounded = np.isfinite(lb) & np.isfinite(ub)
    at_lo = bounded & (x == lb) & (r > 0)
    at_hi = bounded & (x == ub) & (r < 0)
    inner = np.flatnonzero(bounded & ~at_lo & ~at_hi)
    ub[at_lo] = INF
    lb[at_hi] = -INF
    lb[inner[0::2]] = -INF
    ub[inner[1::2]] = INF
    p = QuadraticProgrammingProblem(lb, ub, None, prob.objective_vector, 0.0,
                                    prob.constraint_matrix, prob.right_hand_side,
                                    prob.num_equalities)
    return p, cert


def infeasible_random_lp(n=300, m=150, seed=53):
    """A random LP plus the ">=" row x_a + x_b >= 25 on two variables bounded
    by [0, 10]: primal infeasible."""
    from cupdlp.generators import random_lp
    prob, _ = random_lp(n=n, m=m, seed=seed)
    lb, ub = prob.variable_lower_bound, prob.variable_upper_bound
    a, b = np.flatnonzero(np.isfinite(lb) & np.isfinite(ub))[:2]
    row = sp.csr_matrix(([1.0, 1.0], ([0, 0], [a, b])), shape=(1, n))
    A = sp.vstack([sp.csr_matrix(prob.constraint_matrix), row]).tocsc()
    rhs = np.append(prob.right_hand_side, 25.0)
    return QuadraticProgrammingProblem(lb, ub, None, prob.objective_vector, 0.0, A, rhs,
                                       prob.num_equalities)


def unbounded_random_lp(n=300, m=150, seed=59):
    """A random LP plus a free column with cost -1 and no constraint
    entries: the objective is unbounded below (dual infeasible)."""
    from cupdlp.generators import random_lp
    prob, _ = random_lp(n=n, m=m, seed=seed)
    A = sp.hstack([sp.csc_matrix(prob.constraint_matrix), sp.csc_matrix((m, 1))]).tocsc()
    return QuadraticProgrammingProblem(np.append(prob.variable_lower_bound, -INF),
                                       np.append(prob.variable_upper_bound, INF), None,
                                       np.append(prob.objective_vector, -1.0), 0.0, A,
                                       prob.right_hand_side, prob.num_equalities)


def long_row_lp(n=6000, seed=7):
    """A random LP with one dense row of n entries (longer than a row block)
    plus sparse rows; certified by construction."""
    from cupdlp.generators import random_lp
    prob, cert = random_lp(n=n, m=n // 4, num_eq=n // 8, nnz_per_col=3, seed=seed)
    A = prob.constraint_matrix.tolil()
    A[0, :] = np.random.default_rng(seed).standard_normal(n)
    A = sp.csc_matrix(A)
    b = prob.right_hand_side.copy()
    b[0] = float((A[0, :] @ cert.x)[0])  # equality row 0 stays tight; its y* stays valid
    c = A.T @ cert.y + cert.r
    p = QuadraticProgrammingProblem(prob.variable_lower_bound, prob.variable_upper_bound, None,
                                    c, 0.0, A, b, prob.num_equalities)
    return p, float(c @ cert.x)


def long_col_lp(m=6000, seed=9):
    """A random LP with one dense column of m entries (longer than a slice of
    the primal step) plus sparse columns; certified by construction: b and
    c[0] follow the new column so that the generator's (x*, y*, r) stays
    optimal."""
    from cupdlp.generators import random_lp
    prob, cert = random_lp(n=2 * m, m=m, num_eq=m // 2, nnz_per_col=3, seed=seed)
    A = prob.constraint_matrix.tolil()
    old = np.asarray(A[:, 0].todense()).ravel()
    new = np.random.default_rng(seed).standard_normal(m)
    A[:, 0] = new.reshape(-1, 1)
    A = sp.csc_matrix(A)
    b = prob.right_hand_side + (new - old) * cert.x[0]
    c = prob.objective_vector.copy()
    c[0] = float(new @ cert.y) + cert.r[0]
    p = QuadraticProgrammingProblem(prob.variable_lower_bound, prob.variable_upper_bound, None,
                                    c, 0.0, A, b, prob.num_equalities)
    return p, float(c @ cert.x)


def raw_lp(colptr, rowval, nzval, m, num_eq=0):
    """A pdlp_lp view over hand-written (possibly malformed) 1-based CSC
    arrays; returns (lp, arrays kept alive)."""
    import ctypes as C
    from cupdlp import _abi
    n = len(colptr) - 1
    cp = np.asarray(colptr, dtype=np.int64)
    rv = np.asarray(rowval, dtype=np.int64)
    nz = np.asarray(nzval, dtype=np.float64)
    c, lb, ub, b = np.ones(n), np.zeros(n), np.full(n, 10.0), np.ones(max(m, 1))

    def P(a, t):
        return a.ctypes.data_as(C.POINTER(t))
    lp = _abi.pdlp_lp(n=n, m=m, num_eq=num_eq, nnz=len(rv), colptr=P(cp, C.c_int64),
                      rowval=P(rv, C.c_int64), nzval=P(nz, C.c_double), c=P(c, C.c_double),
                      lb=P(lb, C.c_double), ub=P(ub, C.c_double), b=P(b, C.c_double), c0=0.0)
    return lp, (cp, rv, nz, c, lb, ub, b)
