"""ctypes binding of the CPU oracle (liboracle_pdlp.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and bench.py's
cpu_baseline leg, always as the checker.  It reuses the product's host-side
marshalling (to build the same pdlp_lp / pdlp_params structs) but never
calls the product's solver.
"""
import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "cupdlp.jl_amd"))

from cupdlp import _abi  # noqa: E402
from cupdlp.solver import run_solve, to_c_lp, to_c_params  # noqa: E402

LIB = os.path.join(HERE, "liboracle_pdlp.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_solve.argtypes = [C.POINTER(_abi.pdlp_lp), C.POINTER(_abi.pdlp_params),
                                C.POINTER(_abi.pdlp_result), _abi.STATS_CB, C.c_void_p]
        L.orc_solve.restype = C.c_int
        L.orc_last_error.restype = C.c_char_p
        L.orc_rescale.argtypes = [C.POINTER(_abi.pdlp_lp), C.POINTER(_abi.pdlp_params)] + \
            [C.POINTER(C.c_double)] * 7
        L.orc_rescale.restype = C.c_int
        L.orc_point_stats.argtypes = [C.POINTER(_abi.pdlp_lp), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_double, C.c_double,
                                      C.POINTER(_abi.pdlp_convergence_information),
                                      C.POINTER(_abi.pdlp_infeasibility_information)]
        L.orc_point_stats.restype = C.c_int
        L.orc_step_rule.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64, C.c_double,
                                    C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_step_rule.restype = C.c_double
        L.orc_spmv.argtypes = [C.POINTER(_abi.pdlp_lp), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]
        L.orc_spmv.restype = C.c_int
        L.orc_restart_choice.argtypes = [C.POINTER(_abi.pdlp_restart_params), C.c_int64,
                                         C.c_int64, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.orc_restart_choice.restype = C.c_int32
        L.orc_primal_weight.argtypes = [C.c_double] * 4
        L.orc_primal_weight.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_set_exact.argtypes = [C.c_int]
        L.orc_set_exact.restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def solve(problem, params, want_current=False):
    """The oracle's optimize(params, problem) -> SaddlePointOutput."""
    L = lib()
    rc, out = run_solve(lambda lp, cp, res, cb, u: L.orc_solve(lp, cp, res, cb, u),
                        problem, params, want_current)
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {L.orc_last_error().decode()}")
    return out


def set_exact(on):
    """Exact (compensated) sums for checking; plain FP64 sums for timing."""
    lib().orc_set_exact(1 if on else 0)


def rescale(problem, params):
    """Scaled (nzval, c, l, u, b) and scaling vectors (E, D) of rescale_problem."""
    import numpy as np
    L = lib()
    lp, keep = to_c_lp(problem)
    cp = to_c_params(params)
    n, m, nnz = problem.num_variables, problem.num_constraints, problem.constraint_matrix.nnz
    outs = [np.zeros(max(k, 1)) for k in (nnz, n, n, n, m, m, n)]
    ptrs = [o.ctypes.data_as(C.POINTER(C.c_double)) for o in outs]
    rc = L.orc_rescale(C.byref(lp), C.byref(cp), *ptrs)
    if rc != 0:
        raise OracleError(L.orc_last_error().decode())
    sizes = (nnz, n, n, n, m, m, n)
    return [o[:k] for o, k in zip(outs, sizes)]


def point_stats(problem, x, y, eps_abs=1e-6, eps_rel=1e-6):
    import numpy as np
    L = lib()
    lp, keep = to_c_lp(problem)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ci = _abi.pdlp_convergence_information()
    ii = _abi.pdlp_infeasibility_information()
    rc = L.orc_point_stats(C.byref(lp), x.ctypes.data_as(C.POINTER(C.c_double)),
                           y.ctypes.data_as(C.POINTER(C.c_double)), eps_abs, eps_rel,
                           C.byref(ci), C.byref(ii))
    if rc != 0:
        raise OracleError(L.orc_last_error().decode())
    return ci, ii


def step_rule(eta, interaction, movement, attempt, red=0.3, grow=0.6):
    L = lib()
    acc, ne = C.c_int(), C.c_int()
    v = L.orc_step_rule(eta, interaction, movement, attempt, red, grow, C.byref(acc), C.byref(ne))
    return v, bool(acc.value), bool(ne.value)


def restart_choice(rp, restart_length, iteration, cur_kkt, avg_kkt, last_kkt, last_trial):
    """The oracle's restart decision on given weighted KKT errors: returns
    (RestartChoice, updated last_trial, set of the errors consulted, of
    "cur", "avg", "last")."""
    L = lib()
    crp = to_c_params_restart(rp)
    kkt = (C.c_double * 3)(cur_kkt, avg_kkt, last_kkt)
    lt = C.c_double(last_trial)
    asked = C.c_int32()
    r = L.orc_restart_choice(C.byref(crp), restart_length, iteration, kkt, C.byref(lt),
                             C.byref(asked))
    names = {n for i, n in enumerate(("cur", "avg", "last")) if asked.value >> i & 1}
    return int(r), lt.value, names


def primal_weight(primal_distance, dual_distance, omega, smoothing):
    """The oracle's primal-weight update after a restart."""
    return lib().orc_primal_weight(primal_distance, dual_distance, omega, smoothing)


def to_c_params_restart(rp):
    from cupdlp import PdhgParameters
    return to_c_params(PdhgParameters(restart_params=rp)).restart_params


def spmv(problem, x=None, y=None):
    import numpy as np
    L = lib()
    lp, keep = to_c_lp(problem)
    n, m = problem.num_variables, problem.num_constraints
    Ax = np.zeros(max(m, 1)); ATy = np.zeros(max(n, 1))
    xp = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    yp = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    L.orc_spmv(C.byref(lp),
               None if xp is None else xp.ctypes.data_as(C.POINTER(C.c_double)),
               Ax.ctypes.data_as(C.POINTER(C.c_double)),
               ATy.ctypes.data_as(C.POINTER(C.c_double)),
               None if yp is None else yp.ctypes.data_as(C.POINTER(C.c_double)))
    return Ax[:m], ATy[:n]
