// pdlp_mps.cpp — MPS / QPS reading and the reference's standard form.
//
// Replaces qps_reader_to_standard_form (src/quadratic_programming_io.jl:147-186):
// the file is parsed (free or fixed MPS, optionally gzip-compressed) into the
// two-sided form l_row <= A x <= u_row, l <= x <= u, then
// two_sided_rows_to_slacks (:91-127) and transform_to_standard_form (:33-82)
// produce the QuadraticProgrammingProblem layout: equality rows first, ">="
// rows after, "<=" rows negated, two-sided rows turned into equalities with a
// slack column each.  The reference parses with the QPSReader package; the
// parsing rules below are those of the MPS format it documents:
//   * ROWS: the first N row is the objective, further N rows are dropped;
//     E / L / G rows default to a right-hand side of 0;
//   * COLUMNS: entries of one (row, column) pair add up (as sparse() does);
//     MARKER lines (INTORG / INTEND) are skipped (integrality is dropped);
//   * RHS on the objective row gives the objective constant -value;
//   * RANGES: G -> [rhs, rhs + |R|], L -> [rhs - |R|, rhs],
//     E -> [rhs, rhs + R] for R >= 0, [rhs + R, rhs] for R < 0;
//   * BOUNDS (default [0, +inf)): UP (an upper bound below 0 with the lower
//     bound still 0 makes the lower bound -inf), LO, FX, FR, MI, PL, BV
//     ([0, 1]), LI, UI; SC is rejected;
//   * QUADOBJ / QMATRIX entries are counted (the GPU path has no quadratic
//     term); an OBJSENSE section is rejected, as the reference asserts an
//     unset sense.
// Free format splits lines at blanks; fixed format reads the fields at
// columns 2-3, 5-12, 15-22, 25-36, 40-47 and 50-61.
#include <math.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <string>
#include <string_view>
#include <vector>

#include "pdlp_hip.h"

namespace {

using sv = std::string_view;

struct MpsModel {
  std::vector<char> rtype;            // 'E', 'L', 'G' per constraint row
  std::vector<double> rhs, range;     // range NaN: none
  std::vector<double> c, lb, ub;
  double c0 = 0.0;
  std::vector<int64_t> ei, ej;        // constraint entries (row, col), 0-based
  std::vector<double> ev;
  int64_t q_entries = 0;
};

// Open-addressing name table: a slot holds the 64-bit hash, the value and
// the first 8 bytes of the name, so a lookup of a name of up to 8 bytes
// (fixed-format MPS allows no longer) touches one slot; longer names are
// checked against an arena of the names.
struct NameTable {
  struct Slot {
    uint64_t h;
    char head[8];
    int32_t len;  // -1: empty
    int32_t val;
    uint64_t off;
  };
  std::vector<Slot> slots = std::vector<Slot>(1024, Slot{0, {0}, -1, 0, 0});
  std::string arena;
  size_t count = 0;

  static uint64_t hash(sv s) {  // FNV-1a, then a final mix
    uint64_t h = 1469598103934665603ull;
    for (char ch : s) h = (h ^ (unsigned char)ch) * 1099511628211ull;
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    return h ^ (h >> 32);
  }
  bool same(const Slot &e, uint64_t h, sv s) const {
    if (e.h != h || e.len != (int32_t)s.size()) return false;
    if (memcmp(e.head, s.data(), std::min<size_t>(s.size(), 8)) != 0) return false;
    return s.size() <= 8 || memcmp(arena.data() + e.off + 8, s.data() + 8, s.size() - 8) == 0;
  }
  int find(sv s) const {
    const uint64_t h = hash(s);
    const size_t mask = slots.size() - 1;
    for (size_t i = h & mask;; i = (i + 1) & mask) {
      const Slot &e = slots[i];
      if (e.len < 0) return INT32_MIN;
      if (same(e, h, s)) return e.val;
    }
  }
  void put(const Slot &e) {
    const size_t mask = slots.size() - 1;
    size_t i = e.h & mask;
    while (slots[i].len >= 0) i = (i + 1) & mask;
    slots[i] = e;
  }
  void add(sv s, int v) {
    if (2 * (count + 1) > slots.size()) {
      std::vector<Slot> old(2 * slots.size(), Slot{0, {0}, -1, 0, 0});
      old.swap(slots);
      for (const Slot &e : old)
        if (e.len >= 0) put(e);
    }
    Slot e{hash(s), {0}, (int32_t)s.size(), v, (uint64_t)arena.size()};
    memcpy(e.head, s.data(), std::min<size_t>(s.size(), 8));
    arena.append(s.data(), s.size());
    put(e);
    ++count;
  }
};

inline bool blank(char ch) { return ch == ' ' || ch == '\t' || ch == '\r'; }

// Up to 8 blank-separated tokens of a line.
int split_free(sv line, sv *t) {
  int k = 0;
  size_t p = 0;
  while (p < line.size() && k < 8) {
    while (p < line.size() && blank(line[p])) ++p;
    if (p >= line.size()) break;
    size_t q = p;
    while (q < line.size() && !blank(line[q])) ++q;
    t[k++] = line.substr(p, q - p);
    p = q;
  }
  return k;
}

sv trim(sv s) {
  size_t a = 0, b = s.size();
  while (a < b && blank(s[a])) ++a;
  while (b > a && blank(s[b - 1])) --b;
  return s.substr(a, b - a);
}

// Fixed-format fields 1..6; trailing empty fields dropped.
int split_fixed(sv line, sv *t) {
  static const size_t lo[6] = {1, 4, 14, 24, 39, 49}, hi[6] = {3, 12, 22, 36, 47, 61};
  int k = 0;
  for (int f = 0; f < 6; ++f) {
    if (line.size() <= lo[f]) break;
    t[k++] = trim(line.substr(lo[f], std::min(hi[f], line.size()) - lo[f]));
  }
  while (k > 0 && t[k - 1].empty()) --k;
  return k;
}

// A decimal number (correctly rounded, as strtod), "inf" / "infinity" with
// an optional sign.
bool parse_num(sv s, double *v) {
  if (!s.empty() && s[0] == '+') s.remove_prefix(1);
  if (s.empty()) return false;
  const auto r = std::from_chars(s.data(), s.data() + s.size(), *v);
  return r.ec == std::errc() && r.ptr == s.data() + s.size();
}

struct Reader {
  MpsModel M;
  NameTable rows, cols;  // rows: -1 objective, -2 dropped N row, >= 0 constraint
  std::string err;
  int64_t lineno = 0;
  int last_col = -1;
  std::string last_col_name;
  bool fail(const std::string &m) {
    err = "line " + std::to_string(lineno) + ": " + m;
    return false;
  }
  int col(sv name) {
    if (last_col >= 0 && name == last_col_name) return last_col;
    int j = cols.find(name);
    if (j == INT32_MIN) {
      j = (int)M.c.size();
      cols.add(name, j);
      M.c.push_back(0.0);
      M.lb.push_back(0.0);
      M.ub.push_back(INFINITY);
    }
    last_col = j;
    last_col_name.assign(name.data(), name.size());
    return j;
  }
  bool row(sv name, int *i) {
    *i = rows.find(name);
    if (*i == INT32_MIN) return fail("unknown row " + std::string(name));
    return true;
  }
  bool column_entry(int j, sv rname, sv val) {
    int i;
    double v;
    if (!row(rname, &i)) return false;
    if (!parse_num(val, &v)) return fail("bad value " + std::string(val));
    if (i == -1) {
      M.c[j] += v;
    } else if (i >= 0) {
      M.ei.push_back(i);
      M.ej.push_back(j);
      M.ev.push_back(v);
    }
    return true;
  }
  bool rhs_entry(sv rname, sv val, bool range) {
    int i;
    double v;
    if (!row(rname, &i)) return false;
    if (!parse_num(val, &v)) return fail("bad value " + std::string(val));
    if (range) {
      if (i < 0) return fail("a range on the objective row");
      M.range[i] = v;
    } else if (i == -1) {
      M.c0 = -v;
    } else if (i >= 0) {
      M.rhs[i] = v;
    }
    return true;
  }
};

// Calls fn(line) for every line of a (possibly gzip-compressed) file, read in
// large blocks; stops when fn returns false.  Returns false if the file
// cannot be opened or read.
template <typename F>
bool for_each_line(const char *path, F fn, std::string *err) {
  gzFile f = gzopen(path, "rb");
  if (!f) {
    *err = std::string("cannot open ") + path;
    return false;
  }
  (void)gzbuffer(f, 1 << 20);
  std::vector<char> buf(1 << 22);
  size_t have = 0;
  bool go = true, ok = true;
  for (;;) {
    const int got = gzread(f, buf.data() + have, (unsigned)(buf.size() - have));
    if (got < 0) {
      int code = 0;
      *err = std::string("read error: ") + gzerror(f, &code);
      ok = false;
      break;
    }
    have += (size_t)got;
    const bool eof = got == 0;
    size_t p = 0;
    while (go) {
      const char *nl = (const char *)memchr(buf.data() + p, '\n', have - p);
      if (!nl) break;
      go = fn(sv(buf.data() + p, (size_t)(nl - (buf.data() + p))));
      p = (size_t)(nl - buf.data()) + 1;
    }
    if (!go) break;
    if (eof) {
      if (p < have) go = fn(sv(buf.data() + p, have - p));
      break;
    }
    memmove(buf.data(), buf.data() + p, have - p);
    have -= p;
    if (have == buf.size()) buf.resize(2 * buf.size());  // a line longer than the block
  }
  gzclose(f);
  return ok;
}

bool read_model(const char *path, bool fixed, MpsModel *out, std::string *err) {
  Reader R;
  enum Sec { NONE, ROWS, COLUMNS, RHS, RANGES, BOUNDS, QUAD, END } sec = NONE;
  bool have_obj = false, ok = true;
  sv t[8];
  auto on_line = [&](sv line) -> bool {
    ++R.lineno;
    while (!line.empty() && line.back() == '\r') line.remove_suffix(1);
    if (line.empty() || line[0] == '*') return true;
    MpsModel &M = R.M;
    if (!blank(line[0])) {  // section header
      const int k = split_free(line, t);
      if (k == 0) return true;
      const sv h = t[0];
      if (h == "NAME") sec = NONE;
      else if (h == "ROWS") sec = ROWS;
      else if (h == "COLUMNS") sec = COLUMNS;
      else if (h == "RHS") sec = RHS;
      else if (h == "RANGES") sec = RANGES;
      else if (h == "BOUNDS") sec = BOUNDS;
      else if (h == "QUADOBJ" || h == "QMATRIX" || h == "QSECTION") sec = QUAD;
      else if (h == "ENDATA") { sec = END; return false; }
      else if (h == "OBJSENSE" || h == "OBJSENCE") ok = R.fail("objective sense sections are not supported");
      else ok = R.fail("unsupported section " + std::string(h));
      return ok;
    }
    int k = fixed ? split_fixed(line, t) : split_free(line, t);
    if (k == 0) return true;
    sv *u = t;  // fields after a blank first fixed field
    int nu = k;
    if (fixed && sec != ROWS && sec != BOUNDS && t[0].empty()) {
      u = t + 1;
      nu = k - 1;
    }
    switch (sec) {
      case ROWS: {
        if (k < 2) return ok = R.fail("bad ROWS line");
        const sv ty = t[0], name = t[1];
        if (R.rows.find(name) != INT32_MIN) return ok = R.fail("duplicate row " + std::string(name));
        if (ty == "N") {
          R.rows.add(name, have_obj ? -2 : -1);
          have_obj = true;
        } else if (ty == "E" || ty == "L" || ty == "G") {
          R.rows.add(name, (int)M.rtype.size());
          M.rtype.push_back(ty[0]);
          M.rhs.push_back(0.0);
          M.range.push_back(NAN);
        } else {
          return ok = R.fail("bad row type " + std::string(ty));
        }
        return true;
      }
      case COLUMNS: {
        if (nu >= 3 && (u[1] == "'MARKER'" || u[2] == "'MARKER'")) return true;  // INTORG / INTEND
        if (nu != 3 && nu != 5) return ok = R.fail("bad COLUMNS line");
        const int j = R.col(u[0]);
        ok = R.column_entry(j, u[1], u[2]);
        if (ok && nu == 5) ok = R.column_entry(j, u[3], u[4]);
        return ok;
      }
      case RHS:
      case RANGES: {
        // fixed: blank, set name (may be blank), row, value, row, value
        sv *v = u;
        int nv = nu;
        if (fixed) {
          v = t + 2;
          nv = k > 2 ? k - 2 : 0;
        } else if (nv == 3 || nv == 5) {
          ++v;  // set name
          --nv;
        }
        if (nv != 2 && nv != 4) return ok = R.fail("bad RHS / RANGES line");
        ok = R.rhs_entry(v[0], v[1], sec == RANGES);
        if (ok && nv == 4) ok = R.rhs_entry(v[2], v[3], sec == RANGES);
        return ok;
      }
      case BOUNDS: {
        const sv ty = t[0];
        const bool valued = !(ty == "FR" || ty == "MI" || ty == "PL" || ty == "BV");
        sv *v = t + 1;
        int nv = k - 1;
        if (fixed || (valued && nv == 3) || (!valued && nv == 2)) {
          ++v;  // set name
          --nv;
        }
        if (nv < 1 || (valued && nv < 2)) return ok = R.fail("bad BOUNDS line");
        const int j = R.cols.find(v[0]);
        if (j == INT32_MIN) return ok = R.fail("unknown column " + std::string(v[0]));
        double x = 0.0;
        if (valued && !parse_num(v[1], &x)) return ok = R.fail("bad value " + std::string(v[1]));
        if (ty == "UP" || ty == "UI") {
          M.ub[j] = x;
          if (x < 0.0 && M.lb[j] == 0.0) M.lb[j] = -INFINITY;
        } else if (ty == "LO" || ty == "LI") {
          M.lb[j] = x;
        } else if (ty == "FX") {
          M.lb[j] = x;
          M.ub[j] = x;
        } else if (ty == "FR") {
          M.lb[j] = -INFINITY;
          M.ub[j] = INFINITY;
        } else if (ty == "MI") {
          M.lb[j] = -INFINITY;
        } else if (ty == "PL") {
          M.ub[j] = INFINITY;
        } else if (ty == "BV") {
          M.lb[j] = 0.0;
          M.ub[j] = 1.0;
        } else {
          return ok = R.fail("unsupported bound type " + std::string(ty));
        }
        return true;
      }
      case QUAD: {
        double x;
        if (nu != 3 || !parse_num(u[2], &x)) return ok = R.fail("bad QUADOBJ line");
        if (x != 0.0) ++M.q_entries;
        return true;
      }
      default:
        return ok = R.fail("data outside a section");
    }
  };
  if (!for_each_line(path, on_line, err)) return false;
  if (ok && sec != END) ok = R.fail("missing ENDATA");
  if (ok && !have_obj) ok = R.fail("no objective row");
  if (!ok) {
    *err = R.err;
    return false;
  }
  *out = std::move(R.M);
  return true;
}

// The owned standard-form problem behind a pdlp_mps_problem.
struct Owned {
  pdlp_mps_problem pub{};
  std::vector<int64_t> colptr, rowval;
  std::vector<double> nzval, c, lb, ub, b;
};

}  // namespace

extern "C" {

int pdlp_read_mps(const char *path, int32_t fixed_format, pdlp_mps_problem **out,
                  char *err, int64_t err_len) {
  auto set_err = [&](const std::string &m) {
    if (err && err_len > 0) {
      const size_t k = std::min<size_t>(m.size(), (size_t)err_len - 1);
      memcpy(err, m.data(), k);
      err[k] = '\0';
    }
  };
  if (!path || !out) {
    set_err("null argument");
    return 1;
  }
  *out = nullptr;
  MpsModel M;
  std::string e;
  if (!read_model(path, fixed_format != 0, &M, &e)) {
    set_err(e);
    return 2;
  }
  const int64_t m = (int64_t)M.rtype.size(), n0 = (int64_t)M.c.size();
  // two-sided row bounds from type, right-hand side and range
  std::vector<double> lo(m), hi(m);
  for (int64_t i = 0; i < m; ++i) {
    const double r = M.rhs[i], R = M.range[i];
    switch (M.rtype[i]) {
      case 'E': lo[i] = r; hi[i] = r; break;
      case 'L': lo[i] = -INFINITY; hi[i] = r; break;
      default: lo[i] = r; hi[i] = INFINITY; break;
    }
    if (!isnan(R)) {
      if (M.rtype[i] == 'G') hi[i] = r + fabs(R);
      else if (M.rtype[i] == 'L') lo[i] = r - fabs(R);
      else if (R >= 0.0) hi[i] = r + R;
      else lo[i] = r + R;
    }
  }
  // two_sided_rows_to_slacks (quadratic_programming_io.jl:91-127): one slack
  // column per two-sided row, l <= a'x <= u  ->  a'x - s = 0, l <= s <= u
  std::vector<int64_t> slack_row;
  for (int64_t i = 0; i < m; ++i)
    if (isfinite(lo[i]) && isfinite(hi[i]) && lo[i] != hi[i]) slack_row.push_back(i);
  const int64_t ns = (int64_t)slack_row.size(), n = n0 + ns;
  auto *o = new Owned();
  o->lb = M.lb;
  o->ub = M.ub;
  o->c = M.c;
  for (int64_t k = 0; k < ns; ++k) {
    const int64_t i = slack_row[k];
    o->lb.push_back(lo[i]);
    o->ub.push_back(hi[i]);
    o->c.push_back(0.0);
    lo[i] = 0.0;
    hi[i] = 0.0;
  }
  // transform_to_standard_form (:33-82): equality rows first, then the
  // others in order; "<=" rows negated, rhs = l (or -u for "<=")
  std::vector<char> eq(m), leq(m);
  int64_t neq = 0;
  for (int64_t i = 0; i < m; ++i) {
    eq[i] = lo[i] == hi[i];
    const bool geq = !eq[i] && isfinite(lo[i]);
    leq[i] = !eq[i] && isfinite(hi[i]);
    if (!eq[i] && !geq && !leq[i]) {
      delete o;
      set_err("Not all constraints have finite bounds on at least one side.");
      return 3;
    }
    neq += eq[i];
  }
  std::vector<int64_t> new_of_old(m);
  {
    int64_t k = 0;
    for (int64_t i = 0; i < m; ++i)
      if (eq[i]) new_of_old[i] = k++;
    for (int64_t i = 0; i < m; ++i)
      if (!eq[i]) new_of_old[i] = k++;
  }
  o->b.assign(m, 0.0);
  for (int64_t i = 0; i < m; ++i) o->b[new_of_old[i]] = leq[i] ? -hi[i] : lo[i];
  // CSC: a stable counting sort by column, rows sorted within each column
  // (stable), and the entries of one (row, column) pair added in file order
  struct Ent { int64_t i; double v; };
  const size_t ne = M.ev.size() + (size_t)ns;
  std::vector<int64_t> start(n + 1, 0);
  for (size_t k = 0; k < M.ev.size(); ++k) ++start[M.ej[k] + 1];
  for (int64_t k = 0; k < ns; ++k) ++start[n0 + k + 1];
  for (int64_t j = 0; j < n; ++j) start[j + 1] += start[j];
  std::vector<Ent> ents(ne);
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (size_t k = 0; k < M.ev.size(); ++k) {
      const int64_t i = M.ei[k];
      ents[pos[M.ej[k]]++] = {new_of_old[i], leq[i] ? -M.ev[k] : M.ev[k]};
    }
    for (int64_t k = 0; k < ns; ++k)  // a slack row is an equality now: no sign flip
      ents[pos[n0 + k]++] = {new_of_old[slack_row[k]], -1.0};
  }
  o->colptr.assign(n + 1, 1);
  o->rowval.reserve(ne);
  o->nzval.reserve(ne);
  for (int64_t j = 0; j < n; ++j) {
    Ent *a = ents.data() + start[j], *e = ents.data() + start[j + 1];
    bool sorted = true;
    for (Ent *p = a; p + 1 < e; ++p)
      if (p[1].i < p[0].i) sorted = false;
    if (!sorted) std::stable_sort(a, e, [](const Ent &x, const Ent &y) { return x.i < y.i; });
    for (Ent *p = a; p < e;) {
      Ent *q = p;
      double v = 0.0;
      while (q < e && q->i == p->i) v += (q++)->v;
      o->rowval.push_back(p->i + 1);
      o->nzval.push_back(v);
      p = q;
    }
    o->colptr[j + 1] = (int64_t)o->nzval.size() + 1;
  }
  pdlp_lp &lp = o->pub.lp;
  lp.n = n;
  lp.m = m;
  lp.num_eq = neq;
  lp.nnz = (int64_t)o->nzval.size();
  lp.colptr = o->colptr.data();
  lp.rowval = o->rowval.data();
  lp.nzval = o->nzval.data();
  lp.c = o->c.data();
  lp.lb = o->lb.data();
  lp.ub = o->ub.data();
  lp.b = o->b.data();
  lp.c0 = M.c0;
  o->pub.num_file_variables = n0;
  o->pub.num_slack_variables = ns;
  o->pub.quadratic_entries = M.q_entries;
  o->pub.owner = o;
  *out = &o->pub;
  return 0;
}

void pdlp_mps_free(pdlp_mps_problem *p) {
  if (!p) return;
  delete static_cast<Owned *>(p->owner);
}

}  // extern "C"
