// pdlp_precond.hip — problem preparation on the device: the CSR of A built
// from the CSC input, Ruiz-inf equilibration and Pock-Chambolle rescaling
// (src/preprocess.jl:398-463, 494-525, 541-561), the elementwise matrix norm
// for the initial step size (src/primal_dual_hybrid_gradient_gpu.jl:569-571)
// and the vector norms for the initial primal weight (src/saddle_point_gpu.jl:767-785).
//
// Every value equals the host/oracle computation bit for bit: maxima are
// exact, square roots and divisions are correctly rounded, each entry is
// rescaled as ((1/r_i) * a_ij) * (1/v_j) in both layouts, and the
// Pock-Chambolle sums are compensated sums (order independent).  The row
// kernels walk the slice tables of the step kernels, so a row gets as many
// lanes as its length needs and its entries are read coalesced.  Column
// order within a CSR row follows the scatter's arrival order; the iteration
// kernels' row sums are order independent, so results do not depend on it.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "pdlp_device.h"
#include "pdlp_precond.h"

namespace pdlp {

namespace {

constexpr int kPT = 256;

__device__ __forceinline__ double pjl_max(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a > b) return a;
  if (b > a) return b;
  return __builtin_signbit(a) ? b : a;
}

struct PDD {
  double s, e;
};
__device__ __forceinline__ void pdd_add(PDD &a, double v) {
  const double s = a.s + v;
  const double bb = s - a.s;
  const double t = (a.s - (s - bb)) + (v - bb);
  a.s = s;
  a.e += t;
}
__device__ __forceinline__ void pdd_add_prod(PDD &a, double x, double y) {
  const double p = x * y;
  const double q = __builtin_fma(x, y, -p);
  pdd_add(a, p);
  a.e += q;
}
__device__ __forceinline__ void pdd_merge(PDD &a, const PDD &b) {
  const double s = a.s + b.s;
  const double bb = s - a.s;
  const double t = (a.s - (s - bb)) + (b.s - bb);
  a.s = s;
  a.e = (a.e + b.e) + t;
}

// CSR scatter: one thread per column; arrival order within a row is free.
__global__ __launch_bounds__(kPT) void k_scatter(int n, const int32_t *__restrict__ t_ptr,
                                                 const int32_t *__restrict__ t_col,
                                                 const double *__restrict__ t_val,
                                                 const int32_t *__restrict__ a_ptr,
                                                 int32_t *__restrict__ cursor,
                                                 int32_t *__restrict__ a_col,
                                                 double *__restrict__ a_val) {
  const int j = blockIdx.x * kPT + threadIdx.x;
  if (j >= n) return;
  for (int k = t_ptr[j]; k < t_ptr[j + 1]; ++k) {
    const int i = t_col[k];
    const int p = a_ptr[i] + atomicAdd(&cursor[i], 1);
    a_col[p] = j;
    a_val[p] = t_val[k];
  }
}

// The rows of a slice table (SliceDesc): lane `lane` of a normal slice
// covers entries k = j, j + L, ... of row row0 + g (L = 2^lg lanes per row,
// g = lane / L, j = lane mod L); a chunk slice covers its chunk of one long
// row with all 64 lanes.  `head` marks the lane that owns the row's
// per-row work (the group's first lane; in a chunk slice lane 0 of chunk 0).
struct SliceRow {
  int row, k, step, end;
  bool valid, head;
};

__device__ __forceinline__ SliceRow slice_row(const SliceDesc &sd, int lane,
                                              const int32_t *__restrict__ ptr, bool whole_row) {
  SliceRow r;
  if (sd.nrows == 0) {
    r.row = sd.row0;
    const int len = ptr[r.row + 1] - ptr[r.row];
    r.step = 64;
    r.head = sd.chunk == 0 && lane == 0;
    if (whole_row) {  // reductions: chunk 0's wave takes the whole row
      r.valid = sd.chunk == 0;
      r.k = lane;
      r.end = len;
    } else {
      r.valid = true;
      r.k = sd.chunk * kSellChunk + lane;
      r.end = min(len, (sd.chunk + 1) * kSellChunk);
    }
  } else {
    const int g = lane >> sd.lg, L = 1 << sd.lg;
    r.row = sd.row0 + g;
    r.valid = sd.nrows > 0 && g < sd.nrows;
    r.head = r.valid && (lane & (L - 1)) == 0;
    r.k = lane & (L - 1);
    r.step = L;
    r.end = r.valid ? ptr[r.row + 1] - ptr[r.row] : 0;
  }
  return r;
}

// Lanes of a row: a normal slice's group of 2^lg lanes, a chunk slice's wave.
__device__ __forceinline__ int slice_group_lg(const SliceDesc &sd) {
  return sd.nrows == 0 ? 6 : sd.lg;
}

// sqrt(max |a|) over each row, 0 -> 1 (Ruiz-inf, preprocess.jl:412-454).
__global__ __launch_bounds__(kPT) void k_segmax_sqrt(const SliceDesc *__restrict__ sl, int nsl,
                                                     const int32_t *__restrict__ ptr,
                                                     const double *__restrict__ val,
                                                     double *__restrict__ out) {
  const int wv = (int)blockIdx.x * (kPT / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = sl[wv];
  const SliceRow r = slice_row(sd, lane, ptr, true);
  double mx = 0.0;
  if (r.valid) {
    const double *v = val + ptr[r.row];
    for (int k = r.k; k < r.end; k += r.step) mx = pjl_max(mx, fabs(v[k]));
  }
  for (int o = (1 << slice_group_lg(sd)) >> 1; o > 0; o >>= 1) mx = pjl_max(mx, __shfl_xor(mx, o, 64));
  if (r.valid && r.head) {
    mx = pjl_max(mx, 0.0);  // column: max with the (zero) objective-matrix column
    const double q = sqrt(mx);
    out[r.row] = q == 0.0 ? 1.0 : q;
  }
}

// sqrt(sum |a|^e) over each row (compensated), 0 -> 1; e == 0 counts every
// entry of the dense dimension (0^0 = 1), as Julia's mapreduce does.
__global__ __launch_bounds__(kPT) void k_segpow_sqrt(const SliceDesc *__restrict__ sl, int nsl,
                                                     const int32_t *__restrict__ ptr,
                                                     const double *__restrict__ val, double ex,
                                                     double dense_len, double *__restrict__ out) {
  const int wv = (int)blockIdx.x * (kPT / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = sl[wv];
  const SliceRow r = slice_row(sd, lane, ptr, true);
  PDD acc{0.0, 0.0};
  if (r.valid && ex != 0.0) {
    const double *v = val + ptr[r.row];
    for (int k = r.k; k < r.end; k += r.step) {
      const double a = fabs(v[k]);
      pdd_add(acc, ex == 1.0 ? a : pow(a, ex));
    }
  }
  for (int o = (1 << slice_group_lg(sd)) >> 1; o > 0; o >>= 1)
    pdd_merge(acc, PDD{__shfl_xor(acc.s, o, 64), __shfl_xor(acc.e, o, 64)});
  if (r.valid && r.head) {
    const double q = sqrt(ex == 0.0 ? dense_len : acc.s + acc.e);
    out[r.row] = q == 0.0 ? 1.0 : q;
  }
}

// scale_problem (preprocess.jl:541-561) over the columns: A^T rows, c, l, u, D.
__global__ __launch_bounds__(kPT) void k_scale_cols(const SliceDesc *__restrict__ sl, int nsl,
                                                    const int32_t *__restrict__ t_ptr,
                                                    const int32_t *__restrict__ t_col,
                                                    double *__restrict__ t_val,
                                                    const double *__restrict__ r,
                                                    const double *__restrict__ v,
                                                    double *__restrict__ c, double *__restrict__ l,
                                                    double *__restrict__ u, double *__restrict__ D) {
  const int wv = (int)blockIdx.x * (kPT / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceRow w = slice_row(sl[wv], lane, t_ptr, false);
  if (!w.valid) return;
  const int j = w.row;
  const double vj = v[j];
  if (w.head) {
    c[j] /= vj;
    u[j] *= vj;
    l[j] *= vj;
    D[j] *= vj;
  }
  const double iv = 1.0 / vj;
  const int b = t_ptr[j];
  for (int k = w.k; k < w.end; k += w.step) {
    const double ir = 1.0 / r[t_col[b + k]];
    t_val[b + k] = (ir * t_val[b + k]) * iv;
  }
}

// ... and over the rows: A rows, b, E.
__global__ __launch_bounds__(kPT) void k_scale_rows(const SliceDesc *__restrict__ sl, int nsl,
                                                    const int32_t *__restrict__ a_ptr,
                                                    const int32_t *__restrict__ a_col,
                                                    double *__restrict__ a_val,
                                                    const double *__restrict__ r,
                                                    const double *__restrict__ v,
                                                    double *__restrict__ b, double *__restrict__ E) {
  const int wv = (int)blockIdx.x * (kPT / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceRow w = slice_row(sl[wv], lane, a_ptr, false);
  if (!w.valid) return;
  const int i = w.row;
  const double ri = r[i];
  if (w.head) {
    b[i] /= ri;
    E[i] *= ri;
  }
  const double ir = 1.0 / ri;
  const int p0 = a_ptr[i];
  for (int p = w.k; p < w.end; p += w.step) {
    const double iv = 1.0 / v[a_col[p0 + p]];
    a_val[p0 + p] = (ir * a_val[p0 + p]) * iv;
  }
}

// Block results: [0] max |a| (exact), [1..2] sum b^2, [3..4] sum c^2 (compensated).
__global__ __launch_bounds__(kPT) void k_norms(int64_t nnz, const double *__restrict__ val, int m,
                                               const double *__restrict__ b, int n,
                                               const double *__restrict__ c,
                                               double *__restrict__ part) {
  __shared__ double sm[5][kPT];
  double mx = 0.0;
  PDD sb{0.0, 0.0}, scc{0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kPT;
  for (int64_t k = (int64_t)blockIdx.x * kPT + threadIdx.x; k < nnz; k += stride)
    mx = pjl_max(mx, fabs(val[k]));
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < m; i += stride)
    pdd_add_prod(sb, b[i], b[i]);
  for (int64_t j = (int64_t)blockIdx.x * kPT + threadIdx.x; j < n; j += stride)
    pdd_add_prod(scc, c[j], c[j]);
  sm[0][threadIdx.x] = mx;
  sm[1][threadIdx.x] = sb.s;
  sm[2][threadIdx.x] = sb.e;
  sm[3][threadIdx.x] = scc.s;
  sm[4][threadIdx.x] = scc.e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m0 = sm[0][0];
    PDD a{sm[1][0], sm[2][0]}, cc{sm[3][0], sm[4][0]};
    for (int t = 1; t < kPT; ++t) {
      m0 = pjl_max(m0, sm[0][t]);
      pdd_merge(a, PDD{sm[1][t], sm[2][t]});
      pdd_merge(cc, PDD{sm[3][t], sm[4][t]});
    }
    part[5 * blockIdx.x + 0] = m0;
    part[5 * blockIdx.x + 1] = a.s;
    part[5 * blockIdx.x + 2] = a.e;
    part[5 * blockIdx.x + 3] = cc.s;
    part[5 * blockIdx.x + 4] = cc.e;
  }
}

// Power method of the constant step policy: out = M x over the rows of a
// slice table (whole rows, compensated sums; a chunked row by the wave of
// its first chunk).
__global__ __launch_bounds__(kPT) void k_row_spmv(const SliceDesc *__restrict__ sl, int nsl,
                                                  const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ out) {
  const int wv = (int)blockIdx.x * (kPT / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = sl[wv];
  const SliceRow r = slice_row(sd, lane, ptr, true);
  PDD acc{0.0, 0.0};
  if (r.valid) {
    const int b = ptr[r.row];
    for (int k = r.k; k < r.end; k += r.step) pdd_add_prod(acc, val[b + k], x[col[b + k]]);
  }
  for (int o = (1 << slice_group_lg(sd)) >> 1; o > 0; o >>= 1)
    pdd_merge(acc, PDD{__shfl_xor(acc.s, o, 64), __shfl_xor(acc.e, o, 64)});
  if (r.valid && r.head) out[r.row] = acc.s + acc.e;
}

// Block partials (s, e) of the compensated sum of a[i] * b[i].
__global__ __launch_bounds__(kPT) void k_dot(int64_t len, const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             double *__restrict__ part) {
  __shared__ double sm[2][kPT];
  PDD acc{0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kPT;
  for (int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x; i < len; i += stride)
    pdd_add_prod(acc, a[i], b[i]);
  sm[0][threadIdx.x] = acc.s;
  sm[1][threadIdx.x] = acc.e;
  __syncthreads();
  if (threadIdx.x == 0) {
    PDD t{sm[0][0], sm[1][0]};
    for (int k = 1; k < kPT; ++k) pdd_merge(t, PDD{sm[0][k], sm[1][k]});
    part[2 * blockIdx.x] = t.s;
    part[2 * blockIdx.x + 1] = t.e;
  }
}

__global__ __launch_bounds__(kPT) void k_div(int64_t len, double *__restrict__ x, double d) {
  const int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (i < len) x[i] = x[i] / d;
}

__global__ __launch_bounds__(kPT) void k_to_float(int64_t len, const double *__restrict__ src,
                                                  float *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kPT + threadIdx.x;
  if (i < len) dst[i] = (float)src[i];
}

inline unsigned nb(int64_t len) { return (unsigned)((len + kPT - 1) / kPT); }
inline unsigned nbs(int nsl) { return (unsigned)((nsl + kPT / 64 - 1) / (kPT / 64)); }

}  // namespace

#define PCHK(expr)                     \
  do {                                 \
    hipError_t e_ = (expr);            \
    if (e_ != hipSuccess) return e_;   \
  } while (0)

hipError_t precond_run(hipStream_t s, const PrecondArgs &a, PrecondOut *out) {
  const int n = (int)a.n, m = (int)a.m;
  if (a.nnz > 0 && n > 0)
    hipLaunchKernelGGL(k_scatter, dim3(nb(n)), dim3(kPT), 0, s, n, a.t_ptr, a.t_col, a.t_val,
                       a.a_ptr, a.cursor, a.a_col, a.a_val);
  PCHK(hipGetLastError());
  if (a.after_scatter) PCHK(a.after_scatter(s, a.after_scatter_ctx));
  for (int64_t it = 0; it < a.ruiz_iterations; ++it) {
    if (a.nps > 0)
      hipLaunchKernelGGL(k_segmax_sqrt, dim3(nbs(a.nps)), dim3(kPT), 0, s, a.ps, a.nps, a.t_ptr,
                         a.t_val, a.v);
    if (a.nds > 0)
      hipLaunchKernelGGL(k_segmax_sqrt, dim3(nbs(a.nds)), dim3(kPT), 0, s, a.ds, a.nds, a.a_ptr,
                         a.a_val, a.r);
    if (a.nps > 0)
      hipLaunchKernelGGL(k_scale_cols, dim3(nbs(a.nps)), dim3(kPT), 0, s, a.ps, a.nps, a.t_ptr,
                         a.t_col, a.t_val, a.r, a.v, a.c, a.l, a.u, a.D);
    if (a.nds > 0)
      hipLaunchKernelGGL(k_scale_rows, dim3(nbs(a.nds)), dim3(kPT), 0, s, a.ds, a.nds, a.a_ptr,
                         a.a_col, a.a_val, a.r, a.v, a.b, a.E);
    PCHK(hipGetLastError());
  }
  if (a.has_pc) {
    const double ec = 2.0 - a.pc_alpha, er = a.pc_alpha;
    if (a.nps > 0)
      hipLaunchKernelGGL(k_segpow_sqrt, dim3(nbs(a.nps)), dim3(kPT), 0, s, a.ps, a.nps, a.t_ptr,
                         a.t_val, ec, (double)m, a.v);
    if (a.nds > 0)
      hipLaunchKernelGGL(k_segpow_sqrt, dim3(nbs(a.nds)), dim3(kPT), 0, s, a.ds, a.nds, a.a_ptr,
                         a.a_val, er, (double)n, a.r);
    if (a.nps > 0)
      hipLaunchKernelGGL(k_scale_cols, dim3(nbs(a.nps)), dim3(kPT), 0, s, a.ps, a.nps, a.t_ptr,
                         a.t_col, a.t_val, a.r, a.v, a.c, a.l, a.u, a.D);
    if (a.nds > 0)
      hipLaunchKernelGGL(k_scale_rows, dim3(nbs(a.nds)), dim3(kPT), 0, s, a.ds, a.nds, a.a_ptr,
                         a.a_col, a.a_val, a.r, a.v, a.b, a.E);
    PCHK(hipGetLastError());
  }
  // norms: max |A~| and ||b~||, ||c~|| with a fixed grid so the combine order is fixed
  const int grid = 256;
  hipLaunchKernelGGL(k_norms, dim3(grid), dim3(kPT), 0, s, a.nnz, a.t_val, m, a.b, n, a.c,
                     a.norm_part);
  PCHK(hipGetLastError());
  PCHK(hipMemcpyAsync(a.norm_host, a.norm_part, sizeof(double) * 5 * grid, hipMemcpyDeviceToHost, s));
  PCHK(hipStreamSynchronize(s));
  double mx = 0.0;
  PDD sb{0.0, 0.0}, scc{0.0, 0.0};
  for (int g = 0; g < grid; ++g) {
    const double *p = a.norm_host + 5 * g;
    mx = (p[0] > mx || p[0] != p[0]) ? p[0] : mx;
    // host-side merge, same arithmetic as pdd_merge
    double s1 = sb.s + p[1], bb = s1 - sb.s, t = (sb.s - (s1 - bb)) + (p[1] - bb);
    sb.s = s1; sb.e = (sb.e + p[2]) + t;
    s1 = scc.s + p[3]; bb = s1 - scc.s; t = (scc.s - (s1 - bb)) + (p[3] - bb);
    scc.s = s1; scc.e = (scc.e + p[4]) + t;
  }
  out->max_abs = mx;
  out->norm_b = sqrt(sb.s + sb.e);
  out->norm_c = sqrt(scc.s + scc.e);
  return hipSuccess;
}

static hipError_t dot_host(hipStream_t s, int64_t len, const double *a, const double *b,
                           double *part, double *part_host, double *out) {
  constexpr int grid = 256;
  hipLaunchKernelGGL(k_dot, dim3(grid), dim3(kPT), 0, s, len, a, b, part);
  PCHK(hipGetLastError());
  PCHK(hipMemcpyAsync(part_host, part, sizeof(double) * 2 * grid, hipMemcpyDeviceToHost, s));
  PCHK(hipStreamSynchronize(s));
  PDD t{0.0, 0.0};
  for (int g = 0; g < grid; ++g) {  // host-side merge, same arithmetic as pdd_merge
    const double s1 = t.s + part_host[2 * g], bb = s1 - t.s;
    const double e1 = (t.s - (s1 - bb)) + (part_host[2 * g] - bb);
    t.s = s1;
    t.e = (t.e + part_host[2 * g + 1]) + e1;
  }
  *out = t.s + t.e;
  return hipSuccess;
}

hipError_t power_method(hipStream_t s, const PowerArgs &a, double *sv) {
  const int64_t n = a.n, m = a.m;
  auto spmv_A = [&](const double *x, double *t) -> hipError_t {
    if (a.nds > 0)
      hipLaunchKernelGGL(k_row_spmv, dim3(nbs(a.nds)), dim3(kPT), 0, s, a.ds, a.nds, a.a_ptr,
                         a.a_col, a.a_val, x, t);
    return hipGetLastError();
  };
  auto spmv_AT = [&](const double *t, double *z) -> hipError_t {
    if (a.nps > 0)
      hipLaunchKernelGGL(k_row_spmv, dim3(nbs(a.nps)), dim3(kPT), 0, s, a.ps, a.nps, a.t_ptr,
                         a.t_col, a.t_val, t, z);
    return hipGetLastError();
  };
  for (int64_t it = 0; it < a.iterations; ++it) {
    double nn = 0.0;
    PCHK(dot_host(s, n, a.x, a.x, a.part, a.part_host, &nn));
    const double nx = sqrt(nn);
    if (n > 0) hipLaunchKernelGGL(k_div, dim3(nb(n)), dim3(kPT), 0, s, n, a.x, nx);
    PCHK(hipGetLastError());
    PCHK(spmv_A(a.x, a.t));
    PCHK(spmv_AT(a.t, a.x));
  }
  PCHK(spmv_A(a.x, a.t));
  PCHK(spmv_AT(a.t, a.z));
  double nn = 0.0, xz = 0.0;
  PCHK(dot_host(s, n, a.x, a.x, a.part, a.part_host, &nn));
  PCHK(dot_host(s, n, a.x, a.z, a.part, a.part_host, &xz));
  const double nx = sqrt(nn);
  *sv = sqrt(xz / (nx * nx));
  (void)m;
  return hipSuccess;
}

hipError_t to_float(hipStream_t s, int64_t len, const double *src, float *dst) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_float, dim3(nb(len)), dim3(kPT), 0, s, len, src, dst);
  return hipGetLastError();
}

}  // namespace pdlp
