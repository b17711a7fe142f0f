// pdlp_prep.cpp — validation, cached norms, host rescaling and CSR layouts.
#include "pdlp_prep.h"

#include "pdlp_hostsum.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>

namespace pdlp {

static inline double jl_max(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}

int validate_lp(const pdlp_lp *lp, std::string *err) {
  char buf[512];
  if (!lp) {
    *err = "null problem";
    return 1;
  }
  if (lp->n < 0 || lp->m < 0 || lp->num_eq < 0 || lp->num_eq > lp->m || lp->nnz < 0) {
    snprintf(buf, sizeof buf, "invalid sizes n=%lld m=%lld num_eq=%lld nnz=%lld",
             (long long)lp->n, (long long)lp->m, (long long)lp->num_eq, (long long)lp->nnz);
    *err = buf;
    return 1;
  }
  if (lp->n >= INT32_MAX || lp->m >= INT32_MAX || lp->nnz >= INT32_MAX) {
    *err = "problem exceeds the Int32 device index range";
    return 1;
  }
  if (!lp->colptr || (lp->n > 0 && (!lp->c || !lp->lb || !lp->ub)) || (lp->m > 0 && !lp->b) ||
      (lp->nnz > 0 && (!lp->rowval || !lp->nzval))) {
    *err = "null array in problem";
    return 1;
  }
  if (lp->colptr[0] != 1 || lp->colptr[lp->n] - 1 != lp->nnz) {
    *err = "colptr inconsistent with nnz";
    return 1;
  }
  // colptr in full first (monotone, within 1..nnz+1), so that the row scan
  // below never reads outside rowval
  for (int64_t j = 0; j < lp->n; ++j) {
    if (lp->colptr[j + 1] < lp->colptr[j] || lp->colptr[j + 1] > lp->nnz + 1) {
      snprintf(buf, sizeof buf, "colptr not monotone or out of range at column %lld",
               (long long)j + 1);
      *err = buf;
      return 1;
    }
  }
  for (int64_t j = 0; j < lp->n; ++j) {
    for (int64_t k = lp->colptr[j] - 1; k < lp->colptr[j + 1] - 1; ++k) {
      if (lp->rowval[k] < 1 || lp->rowval[k] > lp->m ||
          (k > lp->colptr[j] - 1 && lp->rowval[k] <= lp->rowval[k - 1])) {
        snprintf(buf, sizeof buf, "row index out of range or unsorted in column %lld",
                 (long long)j + 1);
        *err = buf;
        return 1;
      }
    }
  }
  int bad = 0;
  for (int64_t j = 0; j < lp->n; ++j) {
    if (lp->lb[j] == INFINITY) bad |= 1;
    if (lp->ub[j] == -INFINITY) bad |= 2;
    if (isnan(lp->lb[j]) || isnan(lp->ub[j])) bad |= 4;
    if (!isfinite(lp->c[j])) bad |= 16;
  }
  for (int64_t i = 0; i < lp->m; ++i)
    if (!isfinite(lp->b[i])) bad |= 8;
  for (int64_t k = 0; k < lp->nnz; ++k)
    if (!isfinite(lp->nzval[k])) bad |= 32;
  if (bad) {
    snprintf(buf, sizeof buf,
             "Error found when validating QuadraticProgrammingProblem (flags %d: 1 lb==Inf, "
             "2 ub==-Inf, 4 NaN bound, 8 non-finite rhs, 16 non-finite objective, 32 "
             "non-finite matrix)",
             bad);
    *err = buf;
    return 1;
  }
  return 0;
}

static double norm2(const std::vector<double> &v) { return dd_norm2(v); }

static double norminf(const std::vector<double> &v) {
  double mx = 0.0;
  for (double x : v) {
    const double t = fabs(x);
    if (t > mx || t != t) mx = t;
  }
  return mx;
}

int load_problem(const pdlp_lp *lp, Prepared *P, std::string *err, bool build_csr,
                 int lane_budget) {
  int rc = validate_lp(lp, err);
  if (rc) return rc;
  const int64_t n = lp->n, m = lp->m, nnz = lp->nnz, neq = lp->num_eq;
  P->n = n; P->m = m; P->neq = neq; P->nnz = nnz; P->c0 = lp->c0;
  // Device numbering: columns by nonzero count (descending), rows by nonzero
  // count (descending) within the equality and the ">=" group, ties in the
  // original order.  The vectors the SpMVs gather from then keep their most
  // used entries together in few cache lines.  Every sum the engine forms is
  // compensated (Dot2: error <= eps |sum| + gamma_n^2 sum |terms|), so the
  // numbering changes no computed value unless a sum cancels to ~n^2 eps of
  // its terms; results are mapped back to the caller's numbering.
  std::vector<int64_t> rcount(m, 0);
  for (int64_t k = 0; k < nnz; ++k) rcount[lp->rowval[k] - 1]++;
  // stable counting sorts by descending count
  auto order = [](int64_t lo, int64_t hi, auto count_of, int32_t *out) {
    int64_t mx = 0;
    for (int64_t i = lo; i < hi; ++i) mx = std::max<int64_t>(mx, count_of(i));
    std::vector<int64_t> start(mx + 2, 0);
    for (int64_t i = lo; i < hi; ++i) start[mx - count_of(i) + 1]++;
    for (int64_t c = 0; c <= mx; ++c) start[c + 1] += start[c];
    for (int64_t i = lo; i < hi; ++i) out[lo + start[mx - count_of(i)]++] = (int32_t)i;
  };
  P->col_perm.resize(n);
  P->row_perm.resize(m);
  order(0, n, [&](int64_t j) { return (int64_t)(lp->colptr[j + 1] - lp->colptr[j]); },
        P->col_perm.data());
  auto rc_of = [&](int64_t i) { return rcount[i]; };
  order(0, neq, rc_of, P->row_perm.data());
  order(neq, m, rc_of, P->row_perm.data());
  std::vector<int32_t> new_row(m);
  for (int64_t i = 0; i < m; ++i) new_row[P->row_perm[i]] = (int32_t)i;
  P->t_ptr.resize(n + 1);
  P->t_col.resize(nnz);
  P->t_val.resize(nnz);
  P->t_ptr[0] = 0;
  for (int64_t jn = 0; jn < n; ++jn) {
    const int64_t j = P->col_perm[jn];
    int32_t q = P->t_ptr[jn];
    for (int64_t k = lp->colptr[j] - 1; k < lp->colptr[j + 1] - 1; ++k, ++q) {
      P->t_col[q] = new_row[lp->rowval[k] - 1];
      P->t_val[q] = lp->nzval[k];
    }
    P->t_ptr[jn + 1] = q;
    // row indices of the column ascending in the new numbering, as the
    // column indices of each row of A are: the k-th gathers of a slice's
    // rows then fall in one band of the gathered vector
    const int32_t c0 = P->t_ptr[jn];
    if (q - c0 <= 32) {
      for (int32_t a = c0 + 1; a < q; ++a) {  // insertion sort
        const int32_t r = P->t_col[a];
        const double v = P->t_val[a];
        int32_t b = a;
        for (; b > c0 && P->t_col[b - 1] > r; --b) {
          P->t_col[b] = P->t_col[b - 1];
          P->t_val[b] = P->t_val[b - 1];
        }
        P->t_col[b] = r;
        P->t_val[b] = v;
      }
    } else {
      std::vector<std::pair<int32_t, double>> e((size_t)(q - c0));
      for (int32_t a = c0; a < q; ++a) e[a - c0] = {P->t_col[a], P->t_val[a]};
      std::sort(e.begin(), e.end(),
                [](const std::pair<int32_t, double> &x, const std::pair<int32_t, double> &y) {
                  return x.first < y.first;
                });
      for (int32_t a = c0; a < q; ++a) {
        P->t_col[a] = e[a - c0].first;
        P->t_val[a] = e[a - c0].second;
      }
    }
  }
  P->oc.resize(n); P->ol.resize(n); P->ou.resize(n); P->ob.resize(m);
  for (int64_t jn = 0; jn < n; ++jn) {
    const int64_t j = P->col_perm[jn];
    P->oc[jn] = lp->c[j];
    P->ol[jn] = lp->lb[j];
    P->ou[jn] = lp->ub[j];
  }
  for (int64_t i = 0; i < m; ++i) P->ob[i] = lp->b[P->row_perm[i]];
  P->c = P->oc; P->l = P->ol; P->u = P->ou; P->b = P->ob;
  P->E.assign(m, 1.0);
  P->D.assign(n, 1.0);
  P->qc.l_inf_c = norminf(P->oc);
  P->qc.l_inf_b = norminf(P->ob);
  P->qc.l2_c = norm2(P->oc);
  P->qc.l2_b = norm2(P->ob);
  // CSR of A by a counting sort (column indices in each row ascending)
  P->a_ptr.assign(m + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) P->a_ptr[P->t_col[k] + 1]++;
  for (int64_t i = 0; i < m; ++i) P->a_ptr[i + 1] += P->a_ptr[i];
  build_slices(P->t_ptr.data(), n, &P->ps, lane_budget);
  build_slices(P->a_ptr.data(), m, &P->ds, lane_budget);
  if (P->ps.slots > INT32_MAX || P->ds.slots > INT32_MAX) {
    if (err) *err = "the constraint matrix is too large for 32-bit slot offsets";
    return 3;
  }
  if (!build_csr) {
    P->max_abs_scaled = norminf(P->t_val);
    return 0;
  }
  P->a_col.resize(nnz);
  P->a_val.resize(nnz);
  P->csr_from_csc.resize(nnz);
  std::vector<int32_t> fill(P->a_ptr.begin(), P->a_ptr.end() - (m >= 0 ? 1 : 0));
  for (int64_t j = 0; j < n; ++j)
    for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) {
      const int32_t i = P->t_col[k];
      const int32_t p = fill[i]++;
      P->a_col[p] = (int32_t)j;
      P->a_val[p] = P->t_val[k];
      P->csr_from_csc[p] = k;
    }
  P->max_abs_scaled = norminf(P->t_val);
  return 0;
}

void build_slices(const int32_t *ptr, int64_t R, SliceTable *out, int lane_budget) {
  const int64_t L = lane_budget > 0 ? lane_budget : kSellWidth;
  out->sl.clear();
  out->nchunks.clear();
  int64_t off = 0, r = 0;
  while (r < R) {
    const int64_t len0 = ptr[r + 1] - ptr[r];
    if (len0 > kSellChunk) {  // a long row: chunk slices of kSellChunk entries
      const int32_t nch = (int32_t)((len0 + kSellChunk - 1) / kSellChunk);
      const int32_t lid = (int32_t)out->nchunks.size();
      out->nchunks.push_back(nch);
      for (int32_t c = 0; c < nch; ++c) {
        const int64_t cl = std::min<int64_t>(kSellChunk, len0 - (int64_t)c * kSellChunk);
        const int32_t w = (int32_t)((cl + 63) / 64);
        out->sl.push_back({0, (int32_t)r, (int32_t)off, w, 6, 0, lid, c, 0});
        off += 64 * (int64_t)w;
      }
      ++r;
      continue;
    }
    int lg = 0;
    while (lg < 6 && len0 > (L << lg)) ++lg;
    const int64_t cap = lg == 6 ? kSellChunk : L << lg, maxrows = 64 >> lg;
    int64_t nr = 0, w = 0;
    while (r + nr < R && nr < maxrows) {
      const int64_t len = ptr[r + nr + 1] - ptr[r + nr];
      if (len > cap) break;
      w = std::max<int64_t>(w, (len + (1 << lg) - 1) >> lg);
      ++nr;
    }
    out->sl.push_back({0, (int32_t)r, (int32_t)off, (int32_t)w, (int16_t)lg, (int16_t)nr, 0, 0, 0});
    off += 64 * w;
    r += nr;
  }
  out->slots = off;
}

// scale_problem — src/preprocess.jl:541-561, with Diag(1/r) * A * Diag(1/v)
// evaluated entrywise as ((1/r_i) * a_ij) * (1/v_j).
static void scale(Prepared *P, const std::vector<double> &r, const std::vector<double> &v) {
  for (int64_t j = 0; j < P->n; ++j) {
    P->c[j] /= v[j];
    P->u[j] *= v[j];
    P->l[j] *= v[j];
  }
  for (int64_t i = 0; i < P->m; ++i) P->b[i] /= r[i];
  for (int64_t j = 0; j < P->n; ++j) {
    const double iv = 1.0 / v[j];
    for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) {
      const double ir = 1.0 / r[P->t_col[k]];
      P->t_val[k] = (ir * P->t_val[k]) * iv;
    }
  }
}

void rescale_host(const pdlp_params *prm, Prepared *P) {
  const int64_t n = P->n, m = P->m;
  std::vector<double> r(m), v(n);
  // ruiz_rescaling(p = Inf) — src/preprocess.jl:398-463 (objective matrix zero)
  for (int64_t it = 0; it < prm->l_inf_ruiz_iterations; ++it) {
    for (int64_t j = 0; j < n; ++j) {
      double mx = 0.0;
      for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) mx = jl_max(mx, fabs(P->t_val[k]));
      v[j] = sqrt(jl_max(mx, 0.0));
      if (v[j] == 0.0) v[j] = 1.0;
    }
    if (m > 0) {
      std::fill(r.begin(), r.end(), 0.0);
      for (int64_t j = 0; j < n; ++j)
        for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k)
          r[P->t_col[k]] = jl_max(r[P->t_col[k]], fabs(P->t_val[k]));
      for (int64_t i = 0; i < m; ++i) {
        r[i] = sqrt(r[i]);
        if (r[i] == 0.0) r[i] = 1.0;
      }
    }
    scale(P, r, v);
    for (int64_t i = 0; i < m; ++i) P->E[i] *= r[i];
    for (int64_t j = 0; j < n; ++j) P->D[j] *= v[j];
  }
  // l2_norm_rescaling — src/preprocess.jl:89-103, 344-358
  if (prm->l2_norm_rescaling) {
    // sums of squares as compensated sums (no cancellation, so the value
    // does not depend on the renumbered order)
    std::vector<double> rsf(m, 0.0), csf(n, 0.0);
    std::vector<HostDD> rs(m), cs(n);
    for (int64_t j = 0; j < n; ++j)
      for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) {
        const double a = fabs(P->t_val[k]);
        rsf[P->t_col[k]] = jl_max(rsf[P->t_col[k]], a);
        csf[j] = jl_max(csf[j], a);
      }
    for (auto &x : rsf) if (x == 0.0) x = 1.0;
    for (auto &x : csf) if (x == 0.0) x = 1.0;
    for (int64_t j = 0; j < n; ++j)
      for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) {
        const int32_t i = P->t_col[k];
        const double tr = (1.0 / rsf[i]) * P->t_val[k];
        const double tc = P->t_val[k] * (1.0 / csf[j]);
        rs[i].add_prod(tr, tr);
        cs[j].add_prod(tc, tc);
      }
    for (int64_t i = 0; i < m; ++i) {
      double nr = rsf[i] * sqrt(rs[i].value());
      if (nr == 0.0) nr = 1.0;
      r[i] = sqrt(nr);
    }
    for (int64_t j = 0; j < n; ++j) {
      double nc = csf[j] * sqrt(cs[j].value());
      if (nc == 0.0) nc = 1.0;
      v[j] = sqrt(nc);
    }
    scale(P, r, v);
    for (int64_t i = 0; i < m; ++i) P->E[i] *= r[i];
    for (int64_t j = 0; j < n; ++j) P->D[j] *= v[j];
  }
  // pock_chambolle_rescaling — src/preprocess.jl:494-525
  if (prm->has_pock_chambolle_alpha) {
    const double alpha = prm->pock_chambolle_alpha;
    const double ec = 2.0 - alpha, er = alpha;
    // column and row sums as compensated sums (order independent)
    for (int64_t j = 0; j < n; ++j) {
      double s = 0.0;
      if (ec == 0.0) {
        s = (double)m;
      } else {
        HostDD acc;
        for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k) acc.add(pow(fabs(P->t_val[k]), ec));
        s = acc.value();
      }
      v[j] = sqrt(s);
    }
    if (er == 0.0) {
      std::fill(r.begin(), r.end(), sqrt((double)n));
    } else {
      std::vector<HostDD> racc(m);
      for (int64_t j = 0; j < n; ++j)
        for (int32_t k = P->t_ptr[j]; k < P->t_ptr[j + 1]; ++k)
          racc[P->t_col[k]].add(pow(fabs(P->t_val[k]), er));
      for (int64_t i = 0; i < m; ++i) r[i] = sqrt(racc[i].value());
    }
    for (auto &x : v) if (x == 0.0) x = 1.0;
    for (auto &x : r) if (x == 0.0) x = 1.0;
    scale(P, r, v);
    for (int64_t i = 0; i < m; ++i) P->E[i] *= r[i];
    for (int64_t j = 0; j < n; ++j) P->D[j] *= v[j];
  }
  for (int64_t p = 0; p < P->nnz; ++p) P->a_val[p] = P->t_val[P->csr_from_csc[p]];
  P->max_abs_scaled = norminf(P->t_val);
}

}  // namespace pdlp
