"""MPS / QPS input (src/quadratic_programming_io.jl).

``qps_reader_to_standard_form`` keeps the reference's name and signature
(:147-186); the file is read and put in standard form by the native reader of
libpdlp_hip.so (``pdlp_read_mps``: two_sided_rows_to_slacks :91-127, then
transform_to_standard_form :33-82), host-only.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _abi
from .types import QuadraticProgrammingProblem


class MpsError(ValueError):
    def __init__(self, code, message):
        super().__init__(f"pdlp_read_mps error {code}: {message}")
        self.code = code


def _copy(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


def qps_reader_to_standard_form(filename, fixed_format=False):
    """Reads an MPS file (free format unless ``fixed_format``; ".gz" files are
    decompressed) into a QuadraticProgrammingProblem in standard form:
    equality rows first, ">=" rows after ("<=" rows negated), one slack
    variable per two-sided row after the file's variables.  A nonzero
    quadratic objective raises, as the GPU path has none."""
    lib = _abi.load()
    out = C.POINTER(_abi.pdlp_mps_problem)()
    err = C.create_string_buffer(512)
    rc = lib.pdlp_read_mps(str(filename).encode(), 1 if fixed_format else 0, C.byref(out), err,
                           len(err))
    if rc != 0:
        raise MpsError(rc, err.value.decode(errors="replace"))
    try:
        p = out.contents
        lp = p.lp
        n, m, nnz = lp.n, lp.m, lp.nnz
        if p.quadratic_entries:
            raise ValueError("quadratic objectives are not supported by the GPU path")
        colptr = _copy(lp.colptr, n + 1, np.int64) - 1
        rowval = _copy(lp.rowval, nnz, np.int64) - 1
        A = sp.csc_matrix((_copy(lp.nzval, nnz, np.float64), rowval, colptr), shape=(m, n))
        return QuadraticProgrammingProblem(_copy(lp.lb, n, np.float64), _copy(lp.ub, n, np.float64),
                                           None, _copy(lp.c, n, np.float64), lp.c0, A,
                                           _copy(lp.b, m, np.float64), lp.num_eq)
    finally:
        lib.pdlp_mps_free(out)
