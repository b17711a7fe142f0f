/*
 * pdlp_oracle.c — CPU FP64 restatement of cuPDLP.jl's restarted PDHG.
 *
 * TEST INFRASTRUCTURE.  This library is the checker: only tests/, the
 * smoke check of __graft_entry__.py and the cpu_baseline leg of bench.py
 * load it.  The GPU product (libpdlp_hip.so) never links, loads or calls it.
 *
 * It restates, line by line, the reference's algorithm (paths relative to the
 * cuPDLP.jl tree):
 *   validate                      src/preprocess.jl:8-74
 *   ruiz_rescaling (p = Inf)      src/preprocess.jl:398-463
 *   l2_norm / l2_norm_rescaling   src/preprocess.jl:89-103, 344-358
 *   pock_chambolle_rescaling      src/preprocess.jl:494-525
 *   scale_problem                 src/preprocess.jl:541-561
 *   rescale_problem               src/preprocess.jl:619-675
 *   cached_quadratic_program_info src/termination.jl:130-137
 *   optimize (driver)             src/primal_dual_hybrid_gradient_gpu.jl:457-798
 *   take_step! (adaptive)         src/primal_dual_hybrid_gradient_gpu.jl:343-413
 *   take_step! (constant)         src/primal_dual_hybrid_gradient_gpu.jl:418-452
 *   weighted average              src/saddle_point_gpu.jl:92-240
 *   compute_weight_kkt_residual   src/saddle_point_gpu.jl:250-278
 *   run_restart_scheme & helpers  src/saddle_point_gpu.jl:419-692
 *   select_initial_primal_weight  src/saddle_point_gpu.jl:767-785
 *   KKT statistics                src/iteration_stats_utils_gpu.jl:26-502
 *   check_termination_criteria    src/termination.jl:142-246
 *
 * Choices where the reference leaves the result open (documented):
 *  - Every sum (SpMV rows, dot products, 2-norms, the Pock-Chambolle row and
 *    column sums) is a compensated sum of error-free products rounded once
 *    (Dot2: error <= eps |sum| + gamma_n^2 sum |terms|), i.e. the FP64 value
 *    nearest the exact sum in any order unless the sum cancels to ~n^2 eps of
 *    its terms; cuSPARSE's and cuBLAS's orders are unspecified.  2-norms are
 *    sqrt(sum x^2) without cuBLAS's overflow guard.
 *  - The triple product Diag(1/r) * A * Diag(1/v) of scale_problem is taken
 *    as ((1/r_i) * a_ij) * (1/v_j).
 *  - ConstantStepsizeParams: the power method starts from a SplitMix64/
 *    Box-Muller normal vector (seed 1), not Julia's MersenneTwister(1)
 *    stream, so the estimated step size is not bit-identical to the
 *    reference's; what the reference's estimate satisfies is pinned instead
 *    (0.8 sigma_max <= sigma_hat <= sigma_max against numpy's SVD, and the
 *    iteration count of the power method's stopping rule).
 *
 * Pinning (tests/test_oracle.py, tests/test_oracle_pinning.py): closed-form
 * LPs (optimal, infeasible, unbounded), generator certificates and HiGHS
 * objectives (committed fixtures), Ruiz / Pock-Chambolle / L2 rescaling
 * closed forms and the Pock-Chambolle operator-norm bound, the initial step
 * size 1 / max|A~_ij| and primal weight ||c~|| / ||b~||, KKT statistics at
 * certified optima and under perturbations with closed-form residuals (L2
 * and L_INF), the step-size rule on scripted inputs, the KKT-pass, time and
 * iteration limits and their precedence, and the numerical-error branch
 * (forced movement = 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/pdlp_hip.h"

#define ORC_BLK 4096

static char g_err[1024];

const char *orc_last_error(void) { return g_err; }

static double now_sec(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Julia's min/max on Float64: NaN-propagating, min(-0.0, 0.0) = -0.0,
 * max(-0.0, 0.0) = 0.0. */
static inline double jl_max(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}
static inline double jl_min(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}

/* ---------------- sums ----------------
 * Exact mode (default): every sum is a compensated sum of error-free
 * products (TwoProd with fma, TwoSum), the errors added in FP64 and the
 * total rounded once (Dot2 of Ogita, Rump and Oishi): error <= eps |sum| +
 * gamma_n^2 sum |terms|, so the order changes the result only for sums that
 * cancel to ~n^2 eps of their terms' magnitude.  Plain mode
 * (orc_set_exact(0)) uses ordinary FP64 accumulation and exists only to time
 * an ordinary CPU implementation. */

static int g_exact = 1;

void orc_set_exact(int on) { g_exact = on != 0; }

typedef struct {
  double s, e;
} dd_t;

static inline void dd_add(dd_t *a, double v) {
  double s = a->s + v;
  double bb = s - a->s;
  double t = (a->s - (s - bb)) + (v - bb);
  a->s = s;
  a->e += t;
}
static inline void dd_add_prod(dd_t *a, double x, double y) {
  double p = x * y;
  double q = fma(x, y, -p);
  dd_add(a, p);
  a->e += q;
}
static inline void dd_merge(dd_t *a, dd_t b) {
  double s = a->s + b.s;
  double bb = s - a->s;
  double t = (a->s - (s - bb)) + (b.s - bb);
  a->s = s;
  a->e = (a->e + b.e) + t;
}
static inline double dd_val(dd_t a) { return a.s + a.e; }

/* mode 0: sum a[i]*b[i]; mode 1: sum a[i]; mode 2: sum a[i]^2 + b[i]^2 */
static double blocked(const double *a, const double *b, int64_t len, int mode) {
  if (len <= 0) return 0.0;
  int64_t nb = (len + ORC_BLK - 1) / ORC_BLK;
  dd_t *part = (dd_t *)malloc(sizeof(dd_t) * nb);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nb; ++k) {
    int64_t e = (k + 1) * ORC_BLK < len ? (k + 1) * ORC_BLK : len;
    dd_t acc = {0.0, 0.0};
    if (g_exact) {
      for (int64_t i = k * ORC_BLK; i < e; ++i) {
        if (mode == 0) dd_add_prod(&acc, a[i], b[i]);
        else if (mode == 1) dd_add(&acc, a[i]);
        else { dd_add_prod(&acc, a[i], a[i]); dd_add_prod(&acc, b[i], b[i]); }
      }
    } else {
      double t = 0.0;
      for (int64_t i = k * ORC_BLK; i < e; ++i) {
        if (mode == 0) t += a[i] * b[i];
        else if (mode == 1) t += a[i];
        else t += a[i] * a[i] + b[i] * b[i];
      }
      acc.s = t;
    }
    part[k] = acc;
  }
  dd_t s = {0.0, 0.0};
  for (int64_t k = 0; k < nb; ++k) {
    if (g_exact) dd_merge(&s, part[k]);
    else s.s += part[k].s;
  }
  free(part);
  return g_exact ? dd_val(s) : s.s;
}

static double vdot(const double *a, const double *b, int64_t len) { return blocked(a, b, len, 0); }
static double vsumsq(const double *a, int64_t len) { return blocked(a, a, len, 0); }
static double vsumsq_pair(const double *a, const double *b, int64_t len) { return blocked(a, b, len, 2); }
static double vnorm2(const double *a, int64_t len) { return sqrt(vsumsq(a, len)); }
static double vsum(const double *a, int64_t len) { return blocked(a, a, len, 1); }

/* norm(v, Inf) = maximum(abs, v); 0 for an empty vector. */
static double vmaxabs(const double *a, int64_t len) {
  double mx = 0.0;
  for (int64_t i = 0; i < len; ++i) {
    double t = fabs(a[i]);
    if (t > mx || t != t) mx = t;
  }
  return mx;
}

/* ---------------- problem storage ---------------- */

typedef struct {
  int64_t n, m, neq, nnz;
  int64_t *colptr, *rowidx; /* CSC, 0-based */
  double *cval;
  int64_t *rowptr, *colidx; /* CSR, 0-based */
  double *rval;
  double *c, *l, *u, *b;
  double c0;
} oprob;

static void *xmalloc(size_t sz) {
  void *p = malloc(sz ? sz : 1);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu bytes)\n", sz);
    abort();
  }
  return p;
}

static double *dup_vec(const double *a, int64_t len) {
  double *p = (double *)xmalloc(sizeof(double) * (size_t)(len > 0 ? len : 1));
  if (len > 0) memcpy(p, a, sizeof(double) * (size_t)len);
  return p;
}

static double *zeros(int64_t len) {
  double *p = (double *)calloc((size_t)(len > 0 ? len : 1), sizeof(double));
  if (!p) abort();
  return p;
}

static void oprob_free(oprob *P) {
  free(P->colptr); free(P->rowidx); free(P->cval);
  free(P->rowptr); free(P->colidx); free(P->rval);
  free(P->c); free(P->l); free(P->u); free(P->b);
  memset(P, 0, sizeof(*P));
}

/* CSR from CSC by a counting sort; column indices in a row come out sorted. */
static void build_csr(oprob *P) {
  int64_t m = P->m, n = P->n, nnz = P->nnz;
  free(P->rowptr); free(P->colidx); free(P->rval);
  P->rowptr = (int64_t *)calloc((size_t)m + 1, sizeof(int64_t));
  P->colidx = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
  P->rval = (double *)xmalloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  for (int64_t k = 0; k < nnz; ++k) P->rowptr[P->rowidx[k] + 1]++;
  for (int64_t i = 0; i < m; ++i) P->rowptr[i + 1] += P->rowptr[i];
  int64_t *fill = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
  for (int64_t i = 0; i < m; ++i) fill[i] = P->rowptr[i];
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) {
      int64_t i = P->rowidx[k];
      P->colidx[fill[i]] = j;
      P->rval[fill[i]] = P->cval[k];
      fill[i]++;
    }
  free(fill);
}

/* y = A x over CSR rows, each row summed left to right. */
static void spmv_A(const oprob *P, const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < P->m; ++i) {
    if (g_exact) {
      dd_t acc = {0.0, 0.0};
      for (int64_t k = P->rowptr[i]; k < P->rowptr[i + 1]; ++k)
        dd_add_prod(&acc, P->rval[k], x[P->colidx[k]]);
      y[i] = dd_val(acc);
    } else {
      double s = 0.0;
      for (int64_t k = P->rowptr[i]; k < P->rowptr[i + 1]; ++k) s += P->rval[k] * x[P->colidx[k]];
      y[i] = s;
    }
  }
}

/* y = A^T x over CSC columns (= CSR rows of A^T), each summed top to bottom. */
static void spmv_AT(const oprob *P, const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t j = 0; j < P->n; ++j) {
    if (g_exact) {
      dd_t acc = {0.0, 0.0};
      for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k)
        dd_add_prod(&acc, P->cval[k], x[P->rowidx[k]]);
      y[j] = dd_val(acc);
    } else {
      double s = 0.0;
      for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) s += P->cval[k] * x[P->rowidx[k]];
      y[j] = s;
    }
  }
}

/* ---------------- validation (src/preprocess.jl:8-74) ---------------- */

static int validate_lp(const pdlp_lp *lp) {
  if (lp->n < 0 || lp->m < 0 || lp->num_eq < 0 || lp->num_eq > lp->m || lp->nnz < 0) {
    snprintf(g_err, sizeof g_err, "invalid sizes n=%lld m=%lld num_eq=%lld nnz=%lld",
             (long long)lp->n, (long long)lp->m, (long long)lp->num_eq, (long long)lp->nnz);
    return 1;
  }
  if (lp->n > 0 && (!lp->colptr || !lp->c || !lp->lb || !lp->ub)) {
    snprintf(g_err, sizeof g_err, "null array in problem");
    return 1;
  }
  if (lp->m > 0 && !lp->b) {
    snprintf(g_err, sizeof g_err, "null right_hand_side");
    return 1;
  }
  if (lp->colptr && (lp->colptr[0] != 1 || lp->colptr[lp->n] - 1 != lp->nnz)) {
    snprintf(g_err, sizeof g_err, "colptr inconsistent with nnz");
    return 1;
  }
  if (lp->nnz > 0 && (!lp->rowval || !lp->nzval)) {
    snprintf(g_err, sizeof g_err, "null array in problem");
    return 1;
  }
  /* colptr in full before any rowval read */
  for (int64_t j = 0; j < lp->n; ++j) {
    if (lp->colptr[j + 1] < lp->colptr[j] || lp->colptr[j + 1] > lp->nnz + 1) {
      snprintf(g_err, sizeof g_err, "colptr not monotone or out of range at column %lld",
               (long long)j + 1);
      return 1;
    }
  }
  for (int64_t j = 0; j < lp->n; ++j) {
    for (int64_t k = lp->colptr[j] - 1; k < lp->colptr[j + 1] - 1; ++k) {
      if (lp->rowval[k] < 1 || lp->rowval[k] > lp->m ||
          (k > lp->colptr[j] - 1 && lp->rowval[k] <= lp->rowval[k - 1])) {
        snprintf(g_err, sizeof g_err, "row index out of range or unsorted in column %lld",
                 (long long)j + 1);
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
    snprintf(g_err, sizeof g_err,
             "Error found when validating QuadraticProgrammingProblem (flags %d: 1 lb==Inf, "
             "2 ub==-Inf, 4 NaN bound, 8 non-finite rhs, 16 non-finite objective, 32 "
             "non-finite matrix)",
             bad);
    return 1;
  }
  return 0;
}

static void load_problem(const pdlp_lp *lp, oprob *P) {
  memset(P, 0, sizeof(*P));
  P->n = lp->n; P->m = lp->m; P->neq = lp->num_eq; P->nnz = lp->nnz;
  P->colptr = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(lp->n + 1));
  for (int64_t j = 0; j <= lp->n; ++j) P->colptr[j] = lp->colptr ? lp->colptr[j] - 1 : 0;
  P->rowidx = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(lp->nnz > 0 ? lp->nnz : 1));
  P->cval = (double *)xmalloc(sizeof(double) * (size_t)(lp->nnz > 0 ? lp->nnz : 1));
  for (int64_t k = 0; k < lp->nnz; ++k) {
    P->rowidx[k] = lp->rowval[k] - 1;
    P->cval[k] = lp->nzval[k];
  }
  P->c = dup_vec(lp->c, lp->n);
  P->l = dup_vec(lp->lb, lp->n);
  P->u = dup_vec(lp->ub, lp->n);
  P->b = dup_vec(lp->b, lp->m);
  P->c0 = lp->c0;
}

/* scale_problem — src/preprocess.jl:541-561. */
static void scale_problem(oprob *P, const double *r, const double *v) {
  for (int64_t j = 0; j < P->n; ++j) {
    P->c[j] /= v[j];
    P->u[j] *= v[j];
    P->l[j] *= v[j];
  }
  for (int64_t i = 0; i < P->m; ++i) P->b[i] /= r[i];
  for (int64_t j = 0; j < P->n; ++j) {
    double iv = 1.0 / v[j];
    for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) {
      double ir = 1.0 / r[P->rowidx[k]];
      P->cval[k] = (ir * P->cval[k]) * iv;
    }
  }
}

/* ruiz_rescaling(p = Inf) — src/preprocess.jl:398-463 (objective matrix = 0). */
static void ruiz_inf(oprob *P, int64_t iters, double *cum_r, double *cum_v) {
  double *r = zeros(P->m), *v = zeros(P->n);
  for (int64_t it = 0; it < iters; ++it) {
    for (int64_t j = 0; j < P->n; ++j) {
      double mx = 0.0;
      for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) mx = jl_max(mx, fabs(P->cval[k]));
      mx = jl_max(mx, 0.0); /* max with the (zero) objective-matrix column */
      v[j] = sqrt(mx);
      if (v[j] == 0.0) v[j] = 1.0;
    }
    if (P->m > 0) {
      for (int64_t i = 0; i < P->m; ++i) r[i] = 0.0;
      for (int64_t j = 0; j < P->n; ++j)
        for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) {
          int64_t i = P->rowidx[k];
          r[i] = jl_max(r[i], fabs(P->cval[k]));
        }
      for (int64_t i = 0; i < P->m; ++i) {
        r[i] = sqrt(r[i]);
        if (r[i] == 0.0) r[i] = 1.0;
      }
    }
    scale_problem(P, r, v);
    for (int64_t i = 0; i < P->m; ++i) cum_r[i] *= r[i];
    for (int64_t j = 0; j < P->n; ++j) cum_v[j] *= v[j];
  }
  free(r); free(v);
}

/* l2_norm_rescaling — src/preprocess.jl:89-103, 344-358. */
static void l2_rescale(oprob *P, double *r_out, double *v_out) {
  int64_t m = P->m, n = P->n;
  double *rsf = zeros(m), *csf = zeros(n);
  /* sums of squares as compensated sums (no cancellation: any order gives
     the same value) */
  dd_t *rs = (dd_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(dd_t));
  dd_t *cs = (dd_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(dd_t));
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) {
      int64_t i = P->rowidx[k];
      double a = fabs(P->cval[k]);
      rsf[i] = jl_max(rsf[i], a);
      csf[j] = jl_max(csf[j], a);
    }
  for (int64_t i = 0; i < m; ++i) if (rsf[i] == 0.0) rsf[i] = 1.0;
  for (int64_t j = 0; j < n; ++j) if (csf[j] == 0.0) csf[j] = 1.0;
  /* rows: Diag(1/sf) * A ; columns: A * Diag(1/sf) */
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) {
      int64_t i = P->rowidx[k];
      double tr = (1.0 / rsf[i]) * P->cval[k];
      double tc = P->cval[k] * (1.0 / csf[j]);
      dd_add_prod(&rs[i], tr, tr);
      dd_add_prod(&cs[j], tc, tc);
    }
  for (int64_t i = 0; i < m; ++i) {
    double nr = rsf[i] * sqrt(dd_val(rs[i]));
    if (nr == 0.0) nr = 1.0;
    r_out[i] = sqrt(nr);
  }
  for (int64_t j = 0; j < n; ++j) {
    double nc = csf[j] * sqrt(dd_val(cs[j]));
    if (nc == 0.0) nc = 1.0;
    v_out[j] = sqrt(nc);
  }
  scale_problem(P, r_out, v_out);
  free(rsf); free(csf); free(rs); free(cs);
}

/* pock_chambolle_rescaling — src/preprocess.jl:494-525.  Implicit zeros add
 * abs(0)^e, which is 0 unless e == 0 (then 0^0 = 1 for every entry). */
static void pc_rescale(oprob *P, double alpha, double *r_out, double *v_out) {
  int64_t m = P->m, n = P->n;
  double ec = 2.0 - alpha, er = alpha;
  /* column and row sums as compensated sums (order independent) */
  for (int64_t j = 0; j < n; ++j) {
    double s = 0.0;
    if (ec == 0.0) {
      s = (double)m;
    } else {
      dd_t acc = {0.0, 0.0};
      for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k) dd_add(&acc, pow(fabs(P->cval[k]), ec));
      s = dd_val(acc);
    }
    v_out[j] = sqrt(s);
  }
  if (er == 0.0) {
    for (int64_t i = 0; i < m; ++i) r_out[i] = sqrt((double)n);
  } else {
    dd_t *racc = (dd_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(dd_t));
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = P->colptr[j]; k < P->colptr[j + 1]; ++k)
        dd_add(&racc[P->rowidx[k]], pow(fabs(P->cval[k]), er));
    for (int64_t i = 0; i < m; ++i) r_out[i] = sqrt(dd_val(racc[i]));
    free(racc);
  }
  for (int64_t j = 0; j < n; ++j) if (v_out[j] == 0.0) v_out[j] = 1.0;
  for (int64_t i = 0; i < m; ++i) if (r_out[i] == 0.0) r_out[i] = 1.0;
  scale_problem(P, r_out, v_out);
}

/* rescale_problem — src/preprocess.jl:619-675.  P is a deep copy. */
static void rescale_problem(const pdlp_params *prm, oprob *P, double *E, double *D) {
  for (int64_t i = 0; i < P->m; ++i) E[i] = 1.0;
  for (int64_t j = 0; j < P->n; ++j) D[j] = 1.0;
  if (prm->l_inf_ruiz_iterations > 0) {
    double *cr = zeros(P->m), *cv = zeros(P->n);
    for (int64_t i = 0; i < P->m; ++i) cr[i] = 1.0;
    for (int64_t j = 0; j < P->n; ++j) cv[j] = 1.0;
    ruiz_inf(P, prm->l_inf_ruiz_iterations, cr, cv);
    for (int64_t i = 0; i < P->m; ++i) E[i] *= cr[i];
    for (int64_t j = 0; j < P->n; ++j) D[j] *= cv[j];
    free(cr); free(cv);
  }
  if (prm->l2_norm_rescaling) {
    double *cr = zeros(P->m), *cv = zeros(P->n);
    l2_rescale(P, cr, cv);
    for (int64_t i = 0; i < P->m; ++i) E[i] *= cr[i];
    for (int64_t j = 0; j < P->n; ++j) D[j] *= cv[j];
    free(cr); free(cv);
  }
  if (prm->has_pock_chambolle_alpha) {
    double *cr = zeros(P->m), *cv = zeros(P->n);
    pc_rescale(P, prm->pock_chambolle_alpha, cr, cv);
    for (int64_t i = 0; i < P->m; ++i) E[i] *= cr[i];
    for (int64_t j = 0; j < P->n; ++j) D[j] *= cv[j];
    free(cr); free(cv);
  }
}

/* Exported for tests: scaled problem data and the scaling vectors.
 * Arrays are caller-allocated: cval_out (nnz), c/l/u (n), b (m), E (m), D (n). */
int orc_rescale(const pdlp_lp *lp, const pdlp_params *prm, double *cval_out, double *c_out,
                double *l_out, double *u_out, double *b_out, double *E, double *D) {
  if (validate_lp(lp)) return 1;
  oprob P;
  load_problem(lp, &P);
  rescale_problem(prm, &P, E, D);
  if (P.nnz > 0) memcpy(cval_out, P.cval, sizeof(double) * (size_t)P.nnz);
  if (P.n > 0) {
    memcpy(c_out, P.c, sizeof(double) * (size_t)P.n);
    memcpy(l_out, P.l, sizeof(double) * (size_t)P.n);
    memcpy(u_out, P.u, sizeof(double) * (size_t)P.n);
  }
  if (P.m > 0) memcpy(b_out, P.b, sizeof(double) * (size_t)P.m);
  oprob_free(&P);
  return 0;
}

/* ---------------- KKT statistics ---------------- */

typedef struct {
  double l_inf_c, l_inf_b, l2_c, l2_b;
} qp_cache_t;

/* Buffers of BufferKKTState (src/iteration_stats_utils_gpu.jl:8-20). */
typedef struct {
  double *lv, *uv, *cv, *contrib, *rcv, *rc, *dres;
} kkt_buf;

static void kkt_buf_init(kkt_buf *K, int64_t n, int64_t m, int64_t neq) {
  K->lv = zeros(n); K->uv = zeros(n); K->cv = zeros(m);
  K->contrib = zeros(n); K->rcv = zeros(n); K->rc = zeros(n);
  K->dres = zeros(m - neq);
}
static void kkt_buf_free(kkt_buf *K) {
  free(K->lv); free(K->uv); free(K->cv); free(K->contrib); free(K->rcv); free(K->rc); free(K->dres);
}

/* compute_primal_residual! — src/iteration_stats_utils_gpu.jl:26-94. */
static void primal_residual(int64_t n, int64_t m, int64_t neq, const double *l, const double *u,
                            const double *b, const double *x, const double *Ax, kkt_buf *K) {
  for (int64_t j = 0; j < n; ++j) {
    K->lv[j] = jl_max(l[j] - x[j], 0.0);
    K->uv[j] = jl_max(x[j] - u[j], 0.0);
  }
  for (int64_t i = 0; i < m; ++i) {
    double t = b[i] - Ax[i];
    K->cv[i] = i < neq ? t : jl_max(t, 0.0);
  }
}

static double primal_res_l2(int64_t n, int64_t m, const kkt_buf *K) {
  return sqrt(vsumsq(K->cv, m) + vsumsq_pair(K->lv, K->uv, n));
}
static double primal_res_inf(int64_t n, int64_t m, const kkt_buf *K) {
  return jl_max(jl_max(vmaxabs(K->cv, m), vmaxabs(K->lv, n)), vmaxabs(K->uv, n));
}

/* compute_dual_stats! — src/iteration_stats_utils_gpu.jl:110-238.
 * Returns the dual objective; *dual_res_inf receives norm([dres; rcv], Inf). */
static double dual_stats(int64_t n, int64_t m, int64_t neq, const double *l, const double *u,
                         const double *b, double c0, const double *g, const double *y,
                         kkt_buf *K, double *dual_res_inf) {
  for (int64_t j = 0; j < n; ++j) {
    double fl = isfinite(l[j]) ? 1.0 : 0.0, fu = isfinite(u[j]) ? 1.0 : 0.0;
    K->rc[j] = jl_max(g[j], 0.0) * fl + jl_min(g[j], 0.0) * fu;
    K->rcv[j] = g[j] - K->rc[j];
  }
  for (int64_t i = neq; i < m; ++i) K->dres[i - neq] = jl_max(-y[i], 0.0);
  *dual_res_inf = jl_max(vmaxabs(K->dres, m - neq), vmaxabs(K->rcv, n));
  double base = vdot(b, y, m) + c0;
  for (int64_t j = 0; j < n; ++j) {
    double r = K->rc[j];
    K->contrib[j] = r > 0.0 ? l[j] * r : (r < 0.0 ? u[j] * r : 0.0);
  }
  return base + vsum(K->contrib, n);
}

static double dual_res_l2(int64_t n, int64_t m, int64_t neq, const kkt_buf *K) {
  return sqrt(vsumsq(K->dres, m - neq) + vsumsq(K->rcv, n));
}

/* compute_convergence_information — src/iteration_stats_utils_gpu.jl:251-309. */
static void convergence_info(const oprob *O, const qp_cache_t *qc, const double *x, const double *y,
                             double eps_ratio, int32_t cand, const double *Ax, const double *g,
                             kkt_buf *K, pdlp_convergence_information *ci) {
  int64_t n = O->n, m = O->m, neq = O->neq;
  memset(ci, 0, sizeof(*ci));
  primal_residual(n, m, neq, O->l, O->u, O->b, x, Ax, K);
  ci->primal_objective = O->c0 + vdot(O->c, x, n);
  ci->l_inf_primal_residual = primal_res_inf(n, m, K);
  ci->l2_primal_residual = primal_res_l2(n, m, K);
  ci->relative_l_inf_primal_residual = ci->l_inf_primal_residual / (eps_ratio + qc->l_inf_b);
  ci->relative_l2_primal_residual = ci->l2_primal_residual / (eps_ratio + qc->l2_b);
  ci->l_inf_primal_variable = vmaxabs(x, n);
  ci->l2_primal_variable = vnorm2(x, n);
  double dri;
  ci->dual_objective = dual_stats(n, m, neq, O->l, O->u, O->b, O->c0, g, y, K, &dri);
  ci->l_inf_dual_residual = dri;
  ci->l2_dual_residual = dual_res_l2(n, m, neq, K);
  ci->relative_l_inf_dual_residual = ci->l_inf_dual_residual / (eps_ratio + qc->l_inf_c);
  ci->relative_l2_dual_residual = ci->l2_dual_residual / (eps_ratio + qc->l2_c);
  ci->l_inf_dual_variable = vmaxabs(y, m);
  ci->l2_dual_variable = vnorm2(y, m);
  ci->corrected_dual_objective = dri == 0.0 ? ci->dual_objective : -INFINITY;
  double gap = fabs(ci->primal_objective - ci->dual_objective);
  double abs_obj = fabs(ci->primal_objective) + fabs(ci->dual_objective);
  ci->relative_optimality_gap = gap / (eps_ratio + abs_obj);
  ci->candidate_type = cand;
}

/* compute_infeasibility_information — src/iteration_stats_utils_gpu.jl:314-372.
 * g is the unscaled gradient buffer; it is overwritten with g - c, as the
 * reference's aliased in-place update does (:352). */
static void infeasibility_info(const oprob *O, const double *x, const double *y, int32_t cand,
                               const double *Ax, double *g, kkt_buf *K,
                               pdlp_infeasibility_information *ii) {
  int64_t n = O->n, m = O->m, neq = O->neq;
  memset(ii, 0, sizeof(*ii));
  double s = vmaxabs(x, n);
  double *ray = dup_vec(x, n), *rayAx = dup_vec(Ax, m);
  if (s != 0.0) {
    for (int64_t j = 0; j < n; ++j) ray[j] = x[j] / s;
    for (int64_t i = 0; i < m; ++i) rayAx[i] = Ax[i] / s;
  }
  double *lz = zeros(n), *uz = zeros(n), *bz = zeros(m);
  for (int64_t j = 0; j < n; ++j) {
    double fl = isfinite(O->l[j]) ? 1.0 : 0.0, fu = isfinite(O->u[j]) ? 1.0 : 0.0;
    lz[j] = -1.0 / fl + 1.0;
    uz[j] = 1.0 / fu - 1.0;
  }
  primal_residual(n, m, neq, lz, uz, bz, ray, rayAx, K);
  ii->max_primal_ray_infeasibility = primal_res_inf(n, m, K);
  ii->primal_ray_linear_objective = vdot(O->c, ray, n);
  for (int64_t j = 0; j < n; ++j) g[j] = g[j] - O->c[j];
  double dri;
  double dobj = dual_stats(n, m, neq, O->l, O->u, O->b, O->c0, g, y, K, &dri);
  double scaling = jl_max(vmaxabs(y, m), vmaxabs(K->rc, n));
  if (scaling != 0.0) {
    ii->max_dual_ray_infeasibility = dri / scaling;
    ii->dual_ray_objective = dobj / scaling;
  } else {
    ii->max_dual_ray_infeasibility = 0.0;
    ii->dual_ray_objective = 0.0;
  }
  ii->candidate_type = cand;
  free(ray); free(rayAx); free(lz); free(uz); free(bz);
}

/* compute_weight_kkt_residual — src/saddle_point_gpu.jl:250-278 (scaled space). */
static double weighted_kkt(const oprob *S, const double *x, const double *y, const double *Ax,
                           const double *g, kkt_buf *K, double omega) {
  int64_t n = S->n, m = S->m, neq = S->neq;
  primal_residual(n, m, neq, S->l, S->u, S->b, x, Ax, K);
  double pobj = S->c0 + vdot(S->c, x, n);
  double l2p = primal_res_l2(n, m, K);
  double dri;
  double dobj = dual_stats(n, m, neq, S->l, S->u, S->b, S->c0, g, y, K, &dri);
  double l2d = dual_res_l2(n, m, neq, K);
  double gap = fabs(pobj - dobj);
  return sqrt(omega * (l2p * l2p) + 1.0 / omega * (l2d * l2d) + gap * gap);
}

/* The decision of run_restart_scheme — src/saddle_point_gpu.jl:479-622:
 * the artificial restart (:510-512), the candidate chosen by
 * restart_to_current_metric (should_reset_to_average :419-440), and the
 * adaptive KKT test (should_do_adaptive_restart_kkt :442-474) or the fixed
 * frequency.  kkt(ctx, 0 | 1 | 2) is the weighted KKT error of the current,
 * average or last-restart point, asked for only where the reference computes
 * it.  Returns the RestartChoice; *last_trial is updated on the adaptive
 * branch, *reset_to_average says which point a restart starts from. */
static int32_t restart_choice(const pdlp_restart_params *rp, int64_t restart_length,
                              int64_t iteration, double (*kkt)(void *, int), void *ctx,
                              double *last_trial, int *reset_to_average_out) {
  *reset_to_average_out = 0;
  if (!(restart_length > 0)) return PDLP_RESTART_CHOICE_NO_RESTART;
  int do_restart = 0;
  if ((double)restart_length >= rp->artificial_restart_threshold * (double)(iteration - 1))
    do_restart = 1;
  int reset_to_average = 0;
  double cand = NAN;
  if (rp->restart_scheme != PDLP_NO_RESTARTS) {
    double cur_kkt = kkt(ctx, 0);
    double avg_kkt = kkt(ctx, 1);
    reset_to_average = rp->restart_to_current_metric == PDLP_KKT_GREEDY ? (cur_kkt >= avg_kkt) : 1;
    cand = reset_to_average ? avg_kkt : cur_kkt;
  }
  if (!do_restart) {
    if (rp->restart_scheme == PDLP_ADAPTIVE_KKT) {
      double ratio = cand / kkt(ctx, 2);
      if (ratio < rp->necessary_reduction_for_restart) {
        if (ratio < rp->sufficient_reduction_for_restart)
          do_restart = 1;
        else if (ratio > *last_trial)
          do_restart = 1;
      }
      *last_trial = ratio;
    } else if (rp->restart_scheme == PDLP_FIXED_FREQUENCY &&
               rp->restart_frequency_if_fixed <= restart_length) {
      do_restart = 1;
    }
  }
  *reset_to_average_out = reset_to_average;
  if (!do_restart) return PDLP_RESTART_CHOICE_NO_RESTART;
  return reset_to_average ? PDLP_RESTART_CHOICE_RESTART_TO_AVERAGE
                          : PDLP_RESTART_CHOICE_WEIGHTED_AVERAGE_RESET;
}

/* compute_new_primal_weight — src/saddle_point_gpu.jl:627-654: the
 * log-space moving average of the primal weight toward the ratio of the dual
 * and primal distances moved in the last restart period; unchanged when
 * either distance is at most eps(). */
double orc_primal_weight(double primal_distance, double dual_distance, double omega,
                         double smoothing) {
  const double eps = 2.220446049250313e-16;
  if (!(primal_distance > eps && dual_distance > eps)) return omega;
  const double est = dual_distance / primal_distance;
  const double lw = smoothing * log(est) + (1.0 - smoothing) * log(omega);
  return exp(lw);
}

/* The three weighted KKT errors of the solve's points, for restart_choice. */
typedef struct {
  const oprob *S;
  kkt_buf *K;
  double omega;
  const double *x[3], *y[3], *Ax[3], *g[3];
} kkt_points;

static double kkt_point(void *p, int which) {
  kkt_points *k = (kkt_points *)p;
  return weighted_kkt(k->S, k->x[which], k->y[which], k->Ax[which], k->g[which], k->K, k->omega);
}

/* restart_choice on given errors (tests): kkt3 = current, average,
 * last-restart; *asked gets bit i set when error i was consulted. */
typedef struct {
  const double *v;
  int32_t asked;
} kkt_table;

static double kkt_lookup(void *p, int which) {
  kkt_table *t = (kkt_table *)p;
  t->asked |= 1 << which;
  return t->v[which];
}

int32_t orc_restart_choice(const pdlp_restart_params *rp, int64_t restart_length,
                           int64_t iteration, const double *kkt3, double *last_trial,
                           int32_t *asked) {
  kkt_table t = {kkt3, 0};
  int reset = 0;
  int32_t r = restart_choice(rp, restart_length, iteration, kkt_lookup, &t, last_trial, &reset);
  if (asked) *asked = t.asked;
  return r;
}

/* optimality_criteria_met etc. — src/termination.jl:142-246. */
static int check_termination(const pdlp_termination_criteria *tc, const qp_cache_t *qc,
                             const pdlp_iteration_stats *st) {
  const pdlp_convergence_information *ci = &st->convergence_information;
  double abs_obj = fabs(ci->primal_objective) + fabs(ci->dual_objective);
  double gap = fabs(ci->primal_objective - ci->dual_objective);
  double pe, pb, de, db;
  if (tc->optimality_norm == PDLP_L_INF) {
    pe = ci->l_inf_primal_residual; pb = qc->l_inf_b;
    de = ci->l_inf_dual_residual; db = qc->l_inf_c;
  } else {
    pe = ci->l2_primal_residual; pb = qc->l2_b;
    de = ci->l2_dual_residual; db = qc->l2_c;
  }
  double ea = tc->eps_optimal_absolute, er = tc->eps_optimal_relative;
  if (de < ea + er * db && pe < ea + er * pb && gap < ea + er * abs_obj)
    return PDLP_TERMINATION_REASON_OPTIMAL;
  const pdlp_infeasibility_information *ii = &st->infeasibility_information;
  if (!(ii->dual_ray_objective <= 0.0) &&
      ii->max_dual_ray_infeasibility / ii->dual_ray_objective <= tc->eps_primal_infeasible)
    return PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE;
  if (!(ii->primal_ray_linear_objective >= 0.0) &&
      ii->max_primal_ray_infeasibility / (-ii->primal_ray_linear_objective) <=
          tc->eps_dual_infeasible)
    return PDLP_TERMINATION_REASON_DUAL_INFEASIBLE;
  if ((int64_t)st->iteration_number >= (int64_t)tc->iteration_limit)
    return PDLP_TERMINATION_REASON_ITERATION_LIMIT;
  if (st->cumulative_kkt_matrix_passes >= tc->kkt_matrix_pass_limit)
    return PDLP_TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT;
  if (st->cumulative_time_sec >= tc->time_sec_limit) return PDLP_TERMINATION_REASON_TIME_LIMIT;
  return 0;
}

/* ---------------- power method for the constant step policy ---------------- */

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unif01(uint64_t *s) { return ((double)(splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

/* power_method_failure_probability — src/primal_dual_hybrid_gradient_gpu.jl:139-148 */
static double pm_fail(int64_t dim, double eps, int64_t k) {
  if (k < 2 || eps <= 0.0) return 1.0;
  double a = 0.354 / sqrt(eps * (double)(k - 1));
  return (a < 0.824 ? a : 0.824) * sqrt((double)dim) * pow(1.0 - eps, (double)k - 0.5);
}

/* estimate_maximum_singular_value — src/primal_dual_hybrid_gradient_gpu.jl:150-172 */
static double est_max_sv(const oprob *S, double pfail, double rel_err, int64_t *iters) {
  double eps = 1.0 - (1.0 - rel_err) * (1.0 - rel_err);
  int64_t n = S->n;
  double *x = zeros(n), *t = zeros(S->m);
  uint64_t st = 1;
  for (int64_t j = 0; j < n; ++j) {
    double u1 = unif01(&st), u2 = unif01(&st);
    x[j] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  int64_t k = 0;
  while (pm_fail(n, eps, k) > pfail) {
    double nx = vnorm2(x, n);
    for (int64_t j = 0; j < n; ++j) x[j] /= nx;
    spmv_A(S, x, t);
    spmv_AT(S, t, x);
    k++;
  }
  spmv_A(S, x, t);
  double *z = zeros(n);
  spmv_AT(S, t, z);
  double nx = vnorm2(x, n);
  double sv = sqrt(vdot(x, z, n) / (nx * nx));
  free(x); free(t); free(z);
  *iters = k;
  return sv;
}

/* ---------------- the driver ---------------- */

typedef struct {
  double *x, *y, *Ax, *ATy;
  double *sx, *sy, *sAx, *sATy; /* weighted sums */
  int64_t count;                /* sum_primal_solutions_count (== dual count) */
  double wsum;                  /* sum_primal_solution_weights (== dual) */
  double step_size, primal_weight, kkt_passes;
  int numerical_error;
  int64_t total_iters;
  double *dx, *dy, *Adx;
} ostate;

/* take_step! (adaptive) — src/primal_dual_hybrid_gradient_gpu.jl:343-413.
 * debug_attempt > 0: the attempt with that total_number_iterations sees
 * movement = 0 (fault injection of the numerical-error branch, :384-390). */
static void take_step_adaptive(const oprob *S, ostate *st, double red_exp, double grow_exp,
                               int64_t debug_attempt) {
  int64_t n = S->n, m = S->m, neq = S->neq;
  double eta = st->step_size;
  int done = 0;
  while (!done) {
    st->total_iters += 1;
    double omega = st->primal_weight;
    double tau = eta / omega;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
      double t = st->x[j] - tau * (S->c[j] - st->ATy[j]);
      t = jl_min(S->u[j], jl_max(S->l[j], t));
      st->dx[j] = t - st->x[j];
    }
    spmv_A(S, st->dx, st->Adx);
    double sig = omega * eta;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
      double t = st->y[i] + sig * ((S->b[i] - 2.0 * st->Adx[i]) - st->Ax[i]);
      if (i >= neq) t = jl_max(t, 0.0);
      st->dy[i] = t - st->y[i];
    }
    double interaction = fabs(vdot(st->Adx, st->dy, m));
    double nx = vnorm2(st->dx, n), ny = vnorm2(st->dy, m);
    double movement = 0.5 * omega * (nx * nx) + (0.5 / omega) * (ny * ny);
    if (st->total_iters == debug_attempt) movement = 0.0;
    st->kkt_passes += 1.0;
    double limit;
    if (interaction > 0.0) {
      limit = movement / interaction;
      if (movement == 0.0) {
        st->numerical_error = 1;
        break;
      }
    } else {
      limit = INFINITY;
    }
    if (eta <= limit) {
      double w = st->step_size;
#pragma omp parallel for schedule(static)
      for (int64_t j = 0; j < n; ++j) st->x[j] += st->dx[j];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < m; ++i) {
        st->Ax[i] += st->Adx[i];
        st->y[i] += st->dy[i];
      }
      spmv_AT(S, st->y, st->ATy);
#pragma omp parallel for schedule(static)
      for (int64_t j = 0; j < n; ++j) {
        st->sx[j] += st->x[j] * w;
        st->sATy[j] += st->ATy[j] * w;
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < m; ++i) {
        st->sy[i] += st->y[i] * w;
        st->sAx[i] += st->Ax[i] * w;
      }
      st->count += 1;
      st->wsum += w;
      done = 1;
    }
    double k1 = (double)(st->total_iters + 1);
    double first = (1.0 - 1.0 / pow(k1, red_exp)) * limit;
    double second = (1.0 + 1.0 / pow(k1, grow_exp)) * eta;
    eta = jl_min(first, second);
  }
  st->step_size = eta;
}

/* take_step! (constant) — src/primal_dual_hybrid_gradient_gpu.jl:418-452. */
static void take_step_constant(const oprob *S, ostate *st) {
  int64_t n = S->n, m = S->m, neq = S->neq;
  double eta = st->step_size, omega = st->primal_weight;
  double tau = eta / omega, sig = omega * eta;
  for (int64_t j = 0; j < n; ++j) {
    double t = st->x[j] - tau * (S->c[j] - st->ATy[j]);
    t = jl_min(S->u[j], jl_max(S->l[j], t));
    st->dx[j] = t - st->x[j];
  }
  spmv_A(S, st->dx, st->Adx);
  for (int64_t i = 0; i < m; ++i) {
    double t = st->y[i] + sig * ((S->b[i] - 2.0 * st->Adx[i]) - st->Ax[i]);
    if (i >= neq) t = jl_max(t, 0.0);
    st->dy[i] = t - st->y[i];
  }
  st->kkt_passes += 1.0;
  double w = st->step_size;
  for (int64_t j = 0; j < n; ++j) st->x[j] += st->dx[j];
  for (int64_t i = 0; i < m; ++i) {
    st->Ax[i] += st->Adx[i];
    st->y[i] += st->dy[i];
  }
  spmv_AT(S, st->y, st->ATy);
  for (int64_t j = 0; j < n; ++j) {
    st->sx[j] += st->x[j] * w;
    st->sATy[j] += st->ATy[j] * w;
  }
  for (int64_t i = 0; i < m; ++i) {
    st->sy[i] += st->y[i] * w;
    st->sAx[i] += st->Ax[i] * w;
  }
  st->count += 1;
  st->wsum += w;
}

int orc_solve(const pdlp_lp *lp, const pdlp_params *prm, pdlp_result *res, pdlp_stats_cb cb,
              void *user) {
  double t_call = now_sec();
  g_err[0] = 0;
  if (!lp || !prm || !res) {
    snprintf(g_err, sizeof g_err, "null argument");
    return 1;
  }
  if (validate_lp(lp)) return 1;
  const pdlp_termination_criteria *tc = &prm->termination_criteria;
  const pdlp_restart_params *rp = &prm->restart_params;
  /* construct_restart_parameters asserts — src/saddle_point_gpu.jl:397-403 */
  if (!(rp->restart_frequency_if_fixed > 1) ||
      !(0.0 < rp->artificial_restart_threshold && rp->artificial_restart_threshold <= 1.0) ||
      !(0.0 < rp->sufficient_reduction_for_restart &&
        rp->sufficient_reduction_for_restart <= rp->necessary_reduction_for_restart &&
        rp->necessary_reduction_for_restart <= 1.0) ||
      !(0.0 <= rp->primal_weight_update_smoothing && rp->primal_weight_update_smoothing <= 1.0)) {
    snprintf(g_err, sizeof g_err, "invalid restart parameters");
    return 2;
  }
  if (prm->termination_evaluation_frequency <= 0) {
    snprintf(g_err, sizeof g_err, "termination_evaluation_frequency must be positive");
    return 2;
  }

  oprob O, S;
  load_problem(lp, &O);
  int64_t n = O.n, m = O.m, neq = O.neq;
  qp_cache_t qc = {vmaxabs(O.c, n), vmaxabs(O.b, m), vnorm2(O.c, n), vnorm2(O.b, m)};

  load_problem(lp, &S);
  double *E = zeros(m), *D = zeros(n);
  rescale_problem(prm, &S, E, D);
  build_csr(&S);

  if (!(prm->primal_importance > 0) || !isfinite(prm->primal_importance)) {
    snprintf(g_err, sizeof g_err, "primal_importance must be positive and finite");
    oprob_free(&O); oprob_free(&S); free(E); free(D);
    return 3;
  }

  ostate st;
  memset(&st, 0, sizeof st);
  st.x = zeros(n); st.y = zeros(m); st.Ax = zeros(m); st.ATy = zeros(n);
  st.sx = zeros(n); st.sy = zeros(m); st.sAx = zeros(m); st.sATy = zeros(n);
  st.dx = zeros(n); st.dy = zeros(m); st.Adx = zeros(m);
  st.step_size = 0.0;
  st.primal_weight = 1.0;
  double *avg_x = zeros(n), *avg_y = zeros(m), *avg_Ax = zeros(m), *avg_g = zeros(n);
  double *ox = zeros(n), *oy = zeros(m), *oAx = zeros(m), *og = zeros(n);
  double *gbuf = zeros(n);
  kkt_buf K;
  kkt_buf_init(&K, n, m, neq);
  for (int64_t j = 0; j < n; ++j) gbuf[j] = S.c[j] - st.ATy[j];

  if (prm->step_size_policy == PDLP_STEP_ADAPTIVE) {
    st.kkt_passes += 0.5;
    /* 1 / norm(A, Inf) (elementwise max, pdhg:569-571); a matrix without
       entries (m = 0) would give Inf, where the reference fails in the kernel
       launch: use 1.0 there. */
    double amax = vmaxabs(S.cval, S.nnz);
    st.step_size = amax > 0.0 ? 1.0 / amax : 1.0;
  } else {
    double rel = 0.2;
    int64_t pit = 0;
    double sv = est_max_sv(&S, 0.001, rel, &pit);
    st.step_size = (1.0 - rel) / sv;
    st.kkt_passes += (double)pit;
  }

  if (prm->scale_invariant_initial_primal_weight) {
    double rhs_norm = sqrt(1.0) * vnorm2(S.b, m);
    double obj_norm = sqrt(1.0) * vnorm2(S.c, n);
    if (obj_norm > 0.0 && rhs_norm > 0.0)
      st.primal_weight = prm->primal_importance * (obj_norm / rhs_norm);
    else
      st.primal_weight = prm->primal_importance;
  } else {
    st.primal_weight = prm->primal_importance;
  }
  double smoothing = rp->primal_weight_update_smoothing;

  double start_time = now_sec();
  double basic_time = 0.0;

  /* create_last_restart_info — src/saddle_point_gpu.jl:317-335 */
  double *lx = zeros(n), *ly = zeros(m), *lAx = zeros(m), *lg = dup_vec(gbuf, n);
  double last_pdist = 0.0, last_ddist = 0.0, last_trial = 1.0;

  int64_t iter_limit = tc->iteration_limit;
  int64_t freq = prm->termination_evaluation_frequency;
  st.numerical_error = 0;
  int64_t iteration = 0;
  int rc = 0;
  for (;;) {
    iteration += 1;
    if ((iteration - 1) % freq == 0 || iteration == iter_limit + 1 || iteration <= 10 ||
        st.numerical_error) {
      st.kkt_passes += 2.0;
      if (st.numerical_error || st.count == 0) {
        memcpy(avg_x, st.x, sizeof(double) * n);
        memcpy(avg_y, st.y, sizeof(double) * m);
        memcpy(avg_Ax, st.Ax, sizeof(double) * m);
        memcpy(avg_g, gbuf, sizeof(double) * n);
      } else {
        /* compute_average! — src/saddle_point_gpu.jl:226-240 */
        for (int64_t j = 0; j < n; ++j) {
          avg_x[j] = st.sx[j] / st.wsum;
          avg_g[j] = -st.sATy[j] / st.wsum;
          avg_g[j] += S.c[j];
        }
        for (int64_t i = 0; i < m; ++i) {
          avg_y[i] = st.sy[i] / st.wsum;
          avg_Ax[i] = st.sAx[i] / st.wsum;
        }
      }
      /* evaluate_unscaled_iteration_stats — src/iteration_stats_utils_gpu.jl:447-502 */
      for (int64_t j = 0; j < n; ++j) {
        ox[j] = avg_x[j] / D[j];
        og[j] = avg_g[j] * D[j];
      }
      for (int64_t i = 0; i < m; ++i) {
        oy[i] = avg_y[i] / E[i];
        oAx[i] = avg_Ax[i] * E[i];
      }
      pdlp_iteration_stats stats;
      memset(&stats, 0, sizeof stats);
      stats.iteration_number = (int32_t)(iteration - 1);
      stats.cumulative_kkt_matrix_passes = st.kkt_passes;
      stats.cumulative_time_sec = now_sec() - start_time;
      convergence_info(&O, &qc, ox, oy, tc->eps_optimal_absolute / tc->eps_optimal_relative,
                       PDLP_POINT_TYPE_AVERAGE_ITERATE, oAx, og, &K,
                       &stats.convergence_information);
      infeasibility_info(&O, ox, oy, PDLP_POINT_TYPE_AVERAGE_ITERATE, oAx, og, &K,
                         &stats.infeasibility_information);
      stats.step_size = st.step_size;
      stats.primal_weight = st.primal_weight;
      stats.time_spent_doing_basic_algorithm = basic_time;
      stats.total_attempts = st.total_iters;

      double pnp = 1.0 / st.step_size * st.primal_weight;
      double dnp = 1.0 / st.step_size / st.primal_weight;

      int reason = check_termination(tc, &qc, &stats);
      if (st.numerical_error && reason == 0) reason = PDLP_TERMINATION_REASON_NUMERICAL_ERROR;
      if (iteration < 10 && (reason == PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE ||
                             reason == PDLP_TERMINATION_REASON_DUAL_INFEASIBLE))
        reason = 0;

      if (reason != 0) {
        if (cb) cb(&stats, user);
        for (int64_t j = 0; j < n; ++j) res->x[j] = avg_x[j] / D[j];
        for (int64_t i = 0; i < m; ++i) res->y[i] = avg_y[i] / E[i];
        if (res->x_current)
          for (int64_t j = 0; j < n; ++j) res->x_current[j] = st.x[j] / D[j];
        if (res->y_current)
          for (int64_t i = 0; i < m; ++i) res->y_current[i] = st.y[i] / E[i];
        res->termination_reason = reason;
        res->iteration_count = (int32_t)(iteration - 1);
        res->final_stats = stats;
        res->total_attempts = st.total_iters;
        break;
      }

      for (int64_t j = 0; j < n; ++j) gbuf[j] = S.c[j] - st.ATy[j];

      /* run_restart_scheme — src/saddle_point_gpu.jl:479-622 */
      kkt_points kp = {&S, &K, st.primal_weight,
                       {st.x, avg_x, lx}, {st.y, avg_y, ly}, {st.Ax, avg_Ax, lAx}, {gbuf, avg_g, lg}};
      int reset_to_average = 0;
      const int32_t restart_used = restart_choice(rp, st.count, iteration, kkt_point, &kp,
                                                  &last_trial, &reset_to_average);
      {
        if (restart_used != PDLP_RESTART_CHOICE_NO_RESTART) {
          if (reset_to_average) {
            memcpy(st.x, avg_x, sizeof(double) * n);
            memcpy(st.y, avg_y, sizeof(double) * m);
            memcpy(st.Ax, avg_Ax, sizeof(double) * m);
            for (int64_t j = 0; j < n; ++j) st.ATy[j] = S.c[j] - avg_g[j];
            memcpy(gbuf, avg_g, sizeof(double) * n);
          }
          /* reset_solution_weighted_average! */
          memset(st.sx, 0, sizeof(double) * n);
          memset(st.sy, 0, sizeof(double) * m);
          memset(st.sAx, 0, sizeof(double) * m);
          memset(st.sATy, 0, sizeof(double) * n);
          st.count = 0;
          st.wsum = 0.0;
          /* update_last_restart_info! — src/saddle_point_gpu.jl:659-692 */
          double *dxv = zeros(n), *dyv = zeros(m);
          for (int64_t j = 0; j < n; ++j) dxv[j] = avg_x[j] - lx[j];
          for (int64_t i = 0; i < m; ++i) dyv[i] = avg_y[i] - ly[i];
          last_pdist = sqrt(pnp) * vnorm2(dxv, n) / sqrt(st.primal_weight);
          last_ddist = sqrt(dnp) * vnorm2(dyv, m) * sqrt(st.primal_weight);
          free(dxv); free(dyv);
          memcpy(lx, st.x, sizeof(double) * n);
          memcpy(ly, st.y, sizeof(double) * m);
          memcpy(lAx, st.Ax, sizeof(double) * m);
          memcpy(lg, gbuf, sizeof(double) * n);
        }
      }
      stats.restart_used = restart_used;
      if (restart_used != PDLP_RESTART_CHOICE_NO_RESTART) {
        st.primal_weight = orc_primal_weight(last_pdist, last_ddist, st.primal_weight, smoothing);
      }
      if (cb && prm->record_iteration_stats) cb(&stats, user);
    }
    double t0 = now_sec();
    if (prm->step_size_policy == PDLP_STEP_ADAPTIVE)
      take_step_adaptive(&S, &st, prm->reduction_exponent, prm->growth_exponent,
                         prm->debug_numerical_error_attempt);
    else
      take_step_constant(&S, &st);
    basic_time += now_sec() - t0;
  }
  res->solve_time_sec = now_sec() - t_call;

  free(st.x); free(st.y); free(st.Ax); free(st.ATy);
  free(st.sx); free(st.sy); free(st.sAx); free(st.sATy);
  free(st.dx); free(st.dy); free(st.Adx);
  free(avg_x); free(avg_y); free(avg_Ax); free(avg_g);
  free(ox); free(oy); free(oAx); free(og); free(gbuf);
  free(lx); free(ly); free(lAx); free(lg);
  kkt_buf_free(&K);
  oprob_free(&O); oprob_free(&S);
  free(E); free(D);
  return rc;
}

/* Convergence statistics of a given point in the original space, exactly as
 * the driver evaluates an iterate (unscaled inputs; Ax and the primal
 * gradient c - A^T y are recomputed from the problem data). */
int orc_point_stats(const pdlp_lp *lp, const double *x, const double *y, double eps_abs,
                    double eps_rel, pdlp_convergence_information *ci,
                    pdlp_infeasibility_information *ii) {
  if (validate_lp(lp)) return 1;
  oprob O;
  load_problem(lp, &O);
  build_csr(&O);
  int64_t n = O.n, m = O.m;
  qp_cache_t qc = {vmaxabs(O.c, n), vmaxabs(O.b, m), vnorm2(O.c, n), vnorm2(O.b, m)};
  double *Ax = zeros(m), *ATy = zeros(n), *g = zeros(n);
  spmv_A(&O, x, Ax);
  spmv_AT(&O, y, ATy);
  for (int64_t j = 0; j < n; ++j) g[j] = O.c[j] - ATy[j];
  kkt_buf K;
  kkt_buf_init(&K, n, m, O.neq);
  convergence_info(&O, &qc, x, y, eps_abs / eps_rel, PDLP_POINT_TYPE_CURRENT_ITERATE, Ax, g, &K, ci);
  if (ii) infeasibility_info(&O, x, y, PDLP_POINT_TYPE_CURRENT_ITERATE, Ax, g, &K, ii);
  kkt_buf_free(&K);
  free(Ax); free(ATy); free(g);
  oprob_free(&O);
  return 0;
}

/* One adaptive step-size update of take_step! (src/primal_dual_hybrid_gradient_gpu.jl:384-409)
 * on scripted scalars: returns the new step size; *accepted and *numerical_error
 * report the branch taken. */
double orc_step_rule(double eta, double interaction_signed, double movement, int64_t attempt,
                     double red_exp, double grow_exp, int *accepted, int *numerical_error) {
  double interaction = fabs(interaction_signed);
  double limit;
  *accepted = 0;
  *numerical_error = 0;
  if (interaction > 0.0) {
    limit = movement / interaction;
    if (movement == 0.0) {
      *numerical_error = 1;
      return eta;
    }
  } else {
    limit = INFINITY;
  }
  if (eta <= limit) *accepted = 1;
  double k1 = (double)(attempt + 1);
  double first = (1.0 - 1.0 / pow(k1, red_exp)) * limit;
  double second = (1.0 + 1.0 / pow(k1, grow_exp)) * eta;
  return jl_min(first, second);
}

/* y = A x (CSR order) on the original matrix; exported for tests. */
int orc_spmv(const pdlp_lp *lp, const double *x, double *Ax, double *ATy_of_x_dual,
             const double *ydual) {
  oprob O;
  load_problem(lp, &O);
  build_csr(&O);
  if (x && Ax) spmv_A(&O, x, Ax);
  if (ydual && ATy_of_x_dual) spmv_AT(&O, ydual, ATy_of_x_dual);
  oprob_free(&O);
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
