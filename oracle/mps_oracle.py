"""Test infrastructure only: a pure-Python restatement of MPS reading and of
the reference's standard-form transform, the checker of the native reader
(``pdlp_read_mps``).  Nothing in the product imports it.

Follows src/quadratic_programming_io.jl: ``two_sided_rows_to_slacks``
(:91-127) and ``transform_to_standard_form`` (:33-82) on the two-sided form
the QPSReader package returns (:147-186).  QPSReader itself is not in the
reference tree; the MPS rules restated here (shared with the native reader,
whose header lists them) are the format's: the first N row is the objective
and further N rows are dropped, RHS on the objective row gives the constant
-value, RANGES widen G / L rows by |R| and E rows by R on the side of its
sign, BOUNDS default to [0, +inf) and an UP bound below 0 with the lower bound
still 0 moves the lower bound to -inf, duplicate (row, column) entries add up.

Pinned by the hand-derived standard form of tests/golden/mps/example.mps
(tests/test_mps.py), by a round trip through ``write_free_mps`` of generated
LPs, and by the reference's error on a row with no finite bound.
"""
import gzip
import math

import numpy as np
import scipy.sparse as sp

INF = math.inf
_FIXED_FIELDS = ((1, 3), (4, 12), (14, 22), (24, 36), (39, 47), (49, 61))


class MpsOracleError(ValueError):
    pass


def _fields(line, fixed):
    if not fixed:
        return line.split()
    out = [line[a:b].strip() if len(line) > a else None for a, b in _FIXED_FIELDS]
    out = [f for f in out if f is not None]
    while out and out[-1] == "":
        out.pop()
    return out


def read_two_sided(path, fixed=False):
    """(c, c0, lvar, uvar, lcon, ucon, entries {(i, j): v}, n, m, q_entries)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    rows, rtype, rhs, rng, row_of = [], [], [], [], {}
    cols, col_of, c, lvar, uvar = [], {}, [], [], []
    entries, c0, have_obj, q = {}, 0.0, False, 0
    section = None
    done = False
    with opener(path, "rt") as f:
        for ln, raw in enumerate(f, 1):
            line = raw.rstrip("\r\n")
            if not line or line[0] == "*":
                continue
            if line[0] not in " \t":
                head = line.split()[0]
                if head in ("OBJSENSE", "OBJSENCE"):
                    raise MpsOracleError(f"line {ln}: objective sense sections are not supported")
                if head == "ENDATA":
                    done = True
                    break
                if head not in ("NAME", "ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS",
                                "QUADOBJ", "QMATRIX", "QSECTION"):
                    raise MpsOracleError(f"line {ln}: unsupported section {head}")
                section = head
                continue
            t = _fields(line, fixed)
            if not t:
                continue
            if section == "ROWS":
                ty, name = t[0], t[1]
                if ty == "N":
                    row_of[name] = -2 if have_obj else -1
                    have_obj = True
                else:
                    row_of[name] = len(rows)
                    rows.append(name)
                    rtype.append(ty)
                    rhs.append(0.0)
                    rng.append(None)
            elif section == "COLUMNS":
                u = t[1:] if fixed and t[0] == "" else t
                if "'MARKER'" in u[1:3]:
                    continue
                name = u[0]
                if name not in col_of:
                    col_of[name] = len(cols)
                    cols.append(name)
                    c.append(0.0)
                    lvar.append(0.0)
                    uvar.append(INF)
                j = col_of[name]
                for rname, val in zip(u[1::2], u[2::2]):
                    i = row_of[rname]
                    if i == -1:
                        c[j] += float(val)
                    elif i >= 0:
                        entries[(i, j)] = entries.get((i, j), 0.0) + float(val)
            elif section in ("RHS", "RANGES"):
                if fixed:
                    u = t[2:]
                else:
                    u = t[1:] if len(t) in (3, 5) else t
                for rname, val in zip(u[0::2], u[1::2]):
                    i = row_of[rname]
                    if section == "RANGES":
                        rng[i] = float(val)
                    elif i == -1:
                        c0 = -float(val)
                    elif i >= 0:
                        rhs[i] = float(val)
            elif section == "BOUNDS":
                ty = t[0]
                valued = ty not in ("FR", "MI", "PL", "BV")
                u = t[1:]
                if fixed or (valued and len(u) == 3) or (not valued and len(u) == 2):
                    u = u[1:]
                j = col_of[u[0]]
                v = float(u[1]) if valued else 0.0
                if ty in ("UP", "UI"):
                    uvar[j] = v
                    if v < 0 and lvar[j] == 0.0:
                        lvar[j] = -INF
                elif ty in ("LO", "LI"):
                    lvar[j] = v
                elif ty == "FX":
                    lvar[j] = uvar[j] = v
                elif ty == "FR":
                    lvar[j], uvar[j] = -INF, INF
                elif ty == "MI":
                    lvar[j] = -INF
                elif ty == "PL":
                    uvar[j] = INF
                elif ty == "BV":
                    lvar[j], uvar[j] = 0.0, 1.0
                else:
                    raise MpsOracleError(f"line {ln}: unsupported bound type {ty}")
            else:
                u = t[1:] if fixed and t[0] == "" else t
                if float(u[2]) != 0.0:
                    q += 1
    if not done:
        raise MpsOracleError("missing ENDATA")
    lcon, ucon = [], []
    for ty, r, R in zip(rtype, rhs, rng):
        lo, hi = {"E": (r, r), "L": (-INF, r), "G": (r, INF)}[ty]
        if R is not None:
            if ty == "G":
                hi = r + abs(R)
            elif ty == "L":
                lo = r - abs(R)
            elif R >= 0:
                hi = r + R
            else:
                lo = r + R
        lcon.append(lo)
        ucon.append(hi)
    return c, c0, lvar, uvar, lcon, ucon, entries, len(cols), len(rows), q


def standard_form(path, fixed=False):
    """(l, u, c, c0, A (csc), b, num_eq, num_slacks) as the reference builds
    them from the two-sided form."""
    c, c0, lvar, uvar, lcon, ucon, entries, n0, m, _ = read_two_sided(path, fixed)
    lvar, uvar, c, lcon, ucon = list(lvar), list(uvar), list(c), list(lcon), list(ucon)
    # two_sided_rows_to_slacks (:91-127)
    two = [i for i in range(m) if math.isfinite(lcon[i]) and math.isfinite(ucon[i])
           and lcon[i] != ucon[i]]
    for k, i in enumerate(two):
        lvar.append(lcon[i])
        uvar.append(ucon[i])
        c.append(0.0)
        entries[(i, n0 + k)] = entries.get((i, n0 + k), 0.0) - 1.0
        lcon[i] = ucon[i] = 0.0
    n = n0 + len(two)
    # transform_to_standard_form (:33-82)
    eq = [lcon[i] == ucon[i] for i in range(m)]
    geq = [not eq[i] and math.isfinite(lcon[i]) for i in range(m)]
    leq = [not eq[i] and math.isfinite(ucon[i]) for i in range(m)]
    if sum(eq) + sum(geq) + sum(leq) != m:
        raise MpsOracleError("Not all constraints have finite bounds on at least one side.")
    new_row_to_old = [i for i in range(m) if eq[i]] + [i for i in range(m) if not eq[i]]
    old_to_new = {o: k for k, o in enumerate(new_row_to_old)}
    rows, colsj, vals = [], [], []
    for (i, j), v in entries.items():
        rows.append(old_to_new[i])
        colsj.append(j)
        vals.append(-v if leq[i] else v)
    A = sp.csc_matrix((np.array(vals, dtype=float), (np.array(rows, dtype=np.int64),
                                                     np.array(colsj, dtype=np.int64))),
                      shape=(m, n))
    A.sort_indices()
    b = np.array([(-ucon[i] if leq[i] else lcon[i]) for i in new_row_to_old], dtype=float)
    return (np.array(lvar, dtype=float), np.array(uvar, dtype=float), np.array(c, dtype=float),
            float(c0), A, b, int(sum(eq)), len(two))


def write_free_mps(path, c, c0, l, u, A, b, num_eq):
    """A standard-form LP (equalities, then >= rows) as a free-format MPS file
    whose reading gives the same arrays back (17 significant digits)."""
    A = sp.csc_matrix(A)
    m, n = A.shape
    f = repr
    with open(path, "w") as out:
        out.write("NAME ROUNDTRIP\nROWS\n N OBJ\n")
        for i in range(m):
            out.write(f" {'E' if i < num_eq else 'G'} R{i}\n")
        out.write("COLUMNS\n")
        for j in range(n):
            if c[j] != 0.0:
                out.write(f" X{j} OBJ {f(float(c[j]))}\n")
            for k in range(A.indptr[j], A.indptr[j + 1]):
                out.write(f" X{j} R{A.indices[k]} {f(float(A.data[k]))}\n")
            if c[j] == 0.0 and A.indptr[j] == A.indptr[j + 1]:
                out.write(f" X{j} OBJ 0.0\n")
        out.write("RHS\n")
        if c0 != 0.0:
            out.write(f" RHS OBJ {f(-float(c0))}\n")
        for i in range(m):
            if b[i] != 0.0:
                out.write(f" RHS R{i} {f(float(b[i]))}\n")
        out.write("BOUNDS\n")
        for j in range(n):
            lo, hi = float(l[j]), float(u[j])
            if lo == -INF and hi == INF:
                out.write(f" FR BND X{j}\n")
                continue
            if lo == hi:
                out.write(f" FX BND X{j} {f(lo)}\n")
                continue
            if lo == -INF:
                out.write(f" MI BND X{j}\n")
            elif lo != 0.0:
                out.write(f" LO BND X{j} {f(lo)}\n")
            if hi != INF:
                out.write(f" UP BND X{j} {f(hi)}\n")
        out.write("ENDATA\n")
