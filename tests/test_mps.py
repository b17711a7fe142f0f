"""MPS input: the native reader (pdlp_read_mps, through
cupdlp.qps_reader_to_standard_form) against the hand-derived standard form of
tests/golden/mps/example.mps and against the oracle's restatement of
src/quadratic_programming_io.jl (oracle/mps_oracle.py); fixed format, gzip,
round trips of generated LPs and the error paths.  Host-only: no GPU needed
except for the last test."""
import gzip
import math
import os
import shutil

import numpy as np
import pytest
import scipy.sparse as sp

import mps_oracle as MO
import oracle as O
from cupdlp import MpsError, PdhgParameters, construct_termination_criteria, qps_reader_to_standard_form
from cupdlp.generators import random_lp, transportation_lp

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "golden", "mps", "example.mps")
INF = math.inf

# Standard form of example.mps, derived by hand.  Rows LIM1 (L 4, range 2.5
# -> [1.5, 4]), R4 (G 2, range 3 -> [2, 5]) and R6 (E 3, range -2 -> [1, 3])
# are two-sided: slacks s1, s2, s3 (columns 8-10) with those bounds, and the
# rows become equalities with right-hand side 0.  Row order: the equalities
# LIM1, MYEQN, R4, R6, then LIM2 (>= 1) and R5 (<= 9, negated: >= -9).  The
# second N row (FREE2) is dropped; X4's two R5 entries add up to 4; X4 gets
# MI then UP -2 -> (-inf, -2]; X5's UP -3 with lower bound 0 -> (-inf, -3];
# X6 BV -> [0, 1]; X7 FX 2; RHS -7.5 on the objective -> constant 7.5.
EXP_L = [0.0, -1.0, -INF, -INF, -INF, 0.0, 2.0, 1.5, 2.0, 1.0]
EXP_U = [4.0, 1.0, INF, -2.0, -3.0, 1.0, 2.0, 4.0, 5.0, 3.0]
EXP_C = [1.0, 2.0, -1.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0]
EXP_B = [0.0, 7.0, 0.0, 0.0, 1.0, -9.0]
EXP_A = np.array([
    # X1  X2   X3   X4   X5   X6  X7   s1   s2   s3
    [1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0],   # LIM1
    [0.0, -1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0],   # MYEQN
    [0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0],   # R4
    [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0],   # R6
    [1.0, 0.0, 0.0, 2.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0],   # LIM2
    [0.0, 0.0, 0.0, -4.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0],  # R5 (negated)
])


def check_example(l, u, c, c0, A, b, neq):
    assert list(l) == EXP_L and list(u) == EXP_U and list(c) == EXP_C
    assert c0 == 7.5 and list(b) == EXP_B and neq == 4
    assert np.array_equal(sp.csc_matrix(A).toarray(), EXP_A)


def same_problem(p, q):
    for a, b in ((p.variable_lower_bound, q.variable_lower_bound),
                 (p.variable_upper_bound, q.variable_upper_bound),
                 (p.objective_vector, q.objective_vector), (p.right_hand_side, q.right_hand_side)):
        assert np.array_equal(a, b)
    assert p.objective_constant == q.objective_constant and p.num_equalities == q.num_equalities
    A, B = p.constraint_matrix, q.constraint_matrix
    assert A.shape == B.shape
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data, B.data)


def test_example_native_reader():
    p = qps_reader_to_standard_form(EXAMPLE)
    check_example(p.variable_lower_bound, p.variable_upper_bound, p.objective_vector,
                  p.objective_constant, p.constraint_matrix, p.right_hand_side, p.num_equalities)


def test_example_oracle():
    l, u, c, c0, A, b, neq, nslack = MO.standard_form(EXAMPLE)
    assert nslack == 3
    check_example(l, u, c, c0, A, b, neq)


def fixed_line(f1="", f2="", f3="", f4="", f5="", f6=""):
    buf = [" "] * 61
    for text, start in ((f1, 1), (f2, 4), (f3, 14), (f4, 24), (f5, 39), (f6, 49)):
        for k, ch in enumerate(text):
            buf[start + k] = ch
    return "".join(buf).rstrip()


def write_fixed_example(path):
    """example.mps in fixed format, with X7 renamed "X 7" (fixed-format names
    may hold blanks)."""
    L = ["NAME          EXAMPLE", "ROWS"]
    L += [fixed_line(t, r) for t, r in (("N", "COST"), ("L", "LIM1"), ("G", "LIM2"),
                                        ("E", "MYEQN"), ("G", "R4"), ("L", "R5"), ("E", "R6"),
                                        ("N", "FREE2"))]
    L.append("COLUMNS")
    for col, pairs in (("X1", [("COST", "1.0"), ("LIM1", "1.0")]), ("X1", [("LIM2", "1.0")]),
                       ("X2", [("COST", "2.0"), ("LIM1", "1.0")]), ("X2", [("MYEQN", "-1.0")]),
                       ("X3", [("COST", "-1.0"), ("MYEQN", "1.0")]),
                       ("X3", [("R4", "2.0"), ("FREE2", "5.0")]),
                       ("X4", [("COST", "0.5"), ("R5", "3.0")]),
                       ("X4", [("R5", "1.0"), ("R6", "1.0")]), ("X4", [("LIM2", "2.0")]),
                       ("X5", [("COST", "1.0"), ("LIM2", "-1.0")]), ("X6", [("R5", "1.0")]),
                       ("X 7", [("MYEQN", "1.0")])):
        flat = [x for pr in pairs for x in pr]
        L.append(fixed_line("", col, *flat))
    L.append("RHS")
    for pairs in ([("COST", "-7.5")], [("LIM1", "4.0"), ("LIM2", "1.0")],
                  [("MYEQN", "7.0"), ("R4", "2.0")], [("R5", "9.0"), ("R6", "3.0")]):
        L.append(fixed_line("", "RHS", *[x for pr in pairs for x in pr]))
    L.append("RANGES")
    L.append(fixed_line("", "RNG", "LIM1", "2.5", "R4", "3.0"))
    L.append(fixed_line("", "RNG", "R6", "-2.0"))
    L.append("BOUNDS")
    for ty, col, val in (("UP", "X1", "4.0"), ("LO", "X2", "-1.0"), ("UP", "X2", "1.0"),
                         ("FR", "X3", ""), ("MI", "X4", ""), ("UP", "X4", "-2.0"),
                         ("UP", "X5", "-3.0"), ("BV", "X6", ""), ("FX", "X 7", "2.0")):
        L.append(fixed_line(ty, "BND", col, val))
    L.append("ENDATA")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def test_fixed_format_and_gzip(tmp_path):
    free = qps_reader_to_standard_form(EXAMPLE)
    fx = os.path.join(tmp_path, "example_fixed.mps")
    write_fixed_example(fx)
    same_problem(qps_reader_to_standard_form(fx, fixed_format=True), free)
    l, u, c, c0, A, b, neq, _ = MO.standard_form(fx, fixed=True)
    check_example(l, u, c, c0, A, b, neq)
    gz = os.path.join(tmp_path, "example.mps.gz")
    with open(EXAMPLE, "rb") as fi, gzip.open(gz, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    same_problem(qps_reader_to_standard_form(gz), free)
    check_example(*MO.standard_form(gz)[:7])


@pytest.mark.parametrize("which", ["random", "transport"])
def test_round_trip_generated(tmp_path, which):
    if which == "random":
        prob = random_lp(n=400, m=200, num_eq=80, nnz_per_col=4, seed=5)[0]
    else:
        prob = transportation_lp(12, 15, seed=1)
    prob.objective_constant = 3.25
    path = os.path.join(tmp_path, f"{which}.mps")
    MO.write_free_mps(path, prob.objective_vector, prob.objective_constant,
                      prob.variable_lower_bound, prob.variable_upper_bound,
                      prob.constraint_matrix, prob.right_hand_side, prob.num_equalities)
    got = qps_reader_to_standard_form(path)
    same_problem(got, prob)
    l, u, c, c0, A, b, neq, nslack = MO.standard_form(path)
    assert nslack == 0
    assert np.array_equal(l, got.variable_lower_bound) and np.array_equal(u, got.variable_upper_bound)
    assert np.array_equal(c, got.objective_vector) and np.array_equal(b, got.right_hand_side)
    assert c0 == got.objective_constant and neq == got.num_equalities
    assert (sp.csc_matrix(A) != got.constraint_matrix).nnz == 0


def write(tmp_path, name, text):
    path = os.path.join(tmp_path, name)
    with open(path, "w") as f:
        f.write(text)
    return path


BASE = "NAME T\nROWS\n N OBJ\n {ty} R1\nCOLUMNS\n X1 OBJ 1.0 R1 1.0\nRHS\n RHS R1 {rhs}\n{extra}ENDATA\n"


def test_errors(tmp_path):
    # a row with no finite bound: the reference's error
    p = write(tmp_path, "free_row.mps", BASE.format(ty="G", rhs="-inf", extra=""))
    with pytest.raises(MpsError) as e:
        qps_reader_to_standard_form(p)
    assert e.value.code == 3 and "finite bounds" in str(e.value)
    with pytest.raises(MO.MpsOracleError):
        MO.standard_form(p)
    # an objective sense section (the reference asserts an unset sense)
    p = write(tmp_path, "sense.mps", "NAME T\nOBJSENSE\n    MAX\n" + BASE.format(ty="G", rhs="1", extra="")[7:])
    with pytest.raises(MpsError) as e:
        qps_reader_to_standard_form(p)
    assert e.value.code == 2
    # an entry on an undeclared row; a missing ENDATA; a missing file
    p = write(tmp_path, "bad_row.mps", BASE.format(ty="G", rhs="1", extra="").replace("R1 1.0", "R9 1.0"))
    with pytest.raises(MpsError):
        qps_reader_to_standard_form(p)
    p = write(tmp_path, "no_end.mps", BASE.format(ty="G", rhs="1", extra="").replace("ENDATA\n", ""))
    with pytest.raises(MpsError):
        qps_reader_to_standard_form(p)
    with pytest.raises(MpsError):
        qps_reader_to_standard_form(os.path.join(tmp_path, "absent.mps"))
    # a quadratic objective: read, then refused as the GPU path has none
    p = write(tmp_path, "quad.mps", BASE.format(ty="G", rhs="1", extra="QUADOBJ\n X1 X1 2.0\n"))
    with pytest.raises(ValueError):
        qps_reader_to_standard_form(p)


def test_damaged_files_are_read_or_refused(tmp_path):
    """Damaged copies of the example (truncated, bytes changed, lines dropped,
    repeated or shuffled, tokens replaced by odd ones), free and fixed
    format: each is either read into a consistent problem or refused with
    MpsError / ValueError, never a crash."""
    rng = np.random.default_rng(11)
    lines = open(EXAMPLE).read().split("\n")
    odd = ["nan", "inf", "1e999", "+", "-", "1e", "1.2.3", "", "X" * 300, "R1", "MARKER",
           "'INTORG'", "RANGES", "BOUNDS", "UP", "FR", "BV", "SC", "N", "E", "ENDATA", "\t"]
    read = refused = 0
    for k in range(150):
        t = list(lines)
        op = k % 5
        if op == 0:
            text = "\n".join(t)[: int(rng.integers(0, len("\n".join(t)) + 1))]
        elif op == 1:
            b = bytearray("\n".join(t).encode())
            for _ in range(int(rng.integers(1, 6))):
                b[int(rng.integers(len(b)))] = int(rng.integers(32, 127))
            text = b.decode()
        elif op == 2:
            for _ in range(int(rng.integers(1, 4))):
                del t[int(rng.integers(len(t)))]
            text = "\n".join(t)
        elif op == 3:
            i = int(rng.integers(len(t)))
            t.insert(i, t[i])
            j = int(rng.integers(len(t) - 3))
            t[j:j + 3] = t[j:j + 3][::-1]
            text = "\n".join(t)
        else:
            i = int(rng.integers(len(t)))
            w = t[i].split()
            if w:
                w[int(rng.integers(len(w)))] = odd[int(rng.integers(len(odd)))]
            t[i] = " " + " ".join(w)
            text = "\n".join(t)
        path = os.path.join(tmp_path, f"d{k}.mps")
        with open(path, "w") as f:
            f.write(text)
        for fixed in (False, True):
            try:
                p = qps_reader_to_standard_form(path, fixed_format=fixed)
            except (MpsError, ValueError):
                refused += 1
                continue
            read += 1
            m, n = p.constraint_matrix.shape
            assert p.objective_vector.size == n and p.right_hand_side.size == m
            assert 0 <= p.num_equalities <= m
    assert read > 0 and refused > 0


def test_read_then_oracle_solve(tmp_path):
    """A generated LP written as MPS, read back and solved by the oracle: the
    objective is the generator's certified optimum."""
    prob, cert = random_lp(n=120, m=60, num_eq=30, nnz_per_col=4, seed=8)
    path = os.path.join(tmp_path, "cert.mps")
    MO.write_free_mps(path, prob.objective_vector, 0.0, prob.variable_lower_bound,
                      prob.variable_upper_bound, prob.constraint_matrix, prob.right_hand_side,
                      prob.num_equalities)
    got = qps_reader_to_standard_form(path)
    tc = construct_termination_criteria(eps_optimal_absolute=1e-8, eps_optimal_relative=1e-8)
    out = O.solve(got, PdhgParameters(termination_criteria=tc, record_iteration_stats=False))
    assert out.termination_string == "OPTIMAL"
    obj = float(got.objective_vector @ out.primal_solution)
    assert abs(obj - cert.objective) <= 1e-6 * (1 + abs(cert.objective))


@pytest.mark.gpu
def test_read_then_gpu_solve_matches_in_memory(gpu_ctx, tmp_path):
    print("case: LP (n=2000) written as MPS and read back, GPU solve vs the in-memory LP", flush=True)
    prob = random_lp(n=2000, m=1000, num_eq=500, nnz_per_col=8, seed=3)[0]
    path = os.path.join(tmp_path, "gpu.mps")
    MO.write_free_mps(path, prob.objective_vector, prob.objective_constant,
                      prob.variable_lower_bound, prob.variable_upper_bound,
                      prob.constraint_matrix, prob.right_hand_side, prob.num_equalities)
    got = qps_reader_to_standard_form(path)
    tc = construct_termination_criteria(eps_optimal_absolute=1e-6, eps_optimal_relative=1e-6)
    p = PdhgParameters(termination_criteria=tc, record_iteration_stats=False)
    a, b = gpu_ctx.solve(got, p), gpu_ctx.solve(prob, p)
    assert a.termination_string == "OPTIMAL" and a.iteration_count == b.iteration_count
    assert np.array_equal(a.primal_solution, b.primal_solution)
    assert np.array_equal(a.dual_solution, b.dual_solution)
