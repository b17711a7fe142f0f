// pdlp_kernels.hip — CDNA4 (gfx950) kernels of the restarted-PDHG iteration.
//
// One PDHG step attempt (take_step!, src/primal_dual_hybrid_gradient_gpu.jl:343-413)
// is three launches over a packed set of LP instances:
//   k_primal : per column j of A (= row of A^T):  A^T y gather-SpMV (only when
//              y changed), the pending weighted-sum update of x and A^T y,
//              the projected primal step (K1, :177-197) and sum(dx^2) partials.
//   k_dual   : per row i of A:  A dx gather-SpMV (S1, :225), the pending
//              weighted-sum update of y and Ax, the projected dual step with
//              extrapolation (K2, :232-260) and sum(dy^2), sum(Adx*dy) partials.
//   k_step   : per instance: the fixed-order reduction of those partials and
//              the adaptive step-size rule (:376-409) including accept/reject,
//              with acceptance done by flipping which iterate buffer is current.
// Candidates are written to the non-current buffers, so a rejected attempt
// costs nothing to undo and the host needs no synchronisation per attempt.
//
// Every reduction is deterministic: per-thread sums in a fixed stride order,
// wave butterflies read from lane 0, waves combined in order, and per-block
// partials combined in block order, so results do not depend on timing or on
// which other instances share the launch.
//
// Floating-point contraction is disabled for this file (-ffp-contract=off) so
// each elementwise formula rounds exactly as the reference's written form.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "pdlp_device.h"
#include "pdlp_kernels.h"

namespace pdlp {

// Julia's Float64 max/min: NaN-propagating, max(-0.0, 0.0) = 0.0, min(-0.0, 0.0) = -0.0.
__device__ __forceinline__ double jl_max(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a > b) return a;
  if (b > a) return b;
  return __builtin_signbit(a) ? b : a;
}
__device__ __forceinline__ double jl_min(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a < b) return a;
  if (b < a) return b;
  return __builtin_signbit(a) ? a : b;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = jl_max(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum over the work-group; the result is valid in thread 0.  `scratch`
// holds kThreads/64 doubles.  Contains barriers: call from uniform control flow.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) s += scratch[k];
  }
  return s;
}

template <typename T>
__device__ __forceinline__ double ld(const void *p, int i) {
  return (double)((const T *)p)[i];
}
template <typename T>
__device__ __forceinline__ void st(void *p, int i, double v) {
  ((T *)p)[i] = (T)v;
}

// Products of one row block staged in LDS, then per-row sums by groups of
// `tpr` lanes (a power of two, at most a wave).  epi(row, sum) runs on the
// first lane of each group.  With do_spmv false no matrix entry is read and
// epi receives 0.
template <typename T, typename V, typename Epi>
__device__ __forceinline__ void rowblock_spmv(const int32_t *__restrict__ ptr,
                                              const int32_t *__restrict__ col,
                                              const T *__restrict__ val,
                                              const V *__restrict__ xv, int row0, int nrows,
                                              double *lds, bool do_spmv, Epi &&epi) {
  const int k0 = ptr[row0];
  int tpr = 1;
  if (do_spmv) {
    const int nz = ptr[row0 + nrows] - k0;
    for (int k = threadIdx.x; k < nz; k += kThreads)
      lds[k] = (double)val[k0 + k] * (double)xv[col[k0 + k]];
    __syncthreads();
    while (tpr < 64 && tpr * 2 * nrows <= kThreads) tpr <<= 1;
  }
  const int lane = threadIdx.x & (tpr - 1);
  const int grp = threadIdx.x / tpr;
  const int ngrp = kThreads / tpr;
  for (int r = grp; r < nrows; r += ngrp) {
    double s = 0.0;
    if (do_spmv) {
      const int rs = ptr[row0 + r] - k0, re = ptr[row0 + r + 1] - k0;
      for (int k = rs + lane; k < re; k += tpr) s += lds[k];
      for (int o = tpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, tpr);
    }
    if (lane == 0) epi(row0 + r, s);
  }
}

// Partial sum of one kBlockNnz chunk of a long row; valid in thread 0.
template <typename T, typename V>
__device__ __forceinline__ double longrow_chunk(const int32_t *__restrict__ ptr,
                                                const int32_t *__restrict__ col,
                                                const T *__restrict__ val,
                                                const V *__restrict__ xv, int row, int chunk,
                                                double *scratch) {
  const int kb = ptr[row] + chunk * kBlockNnz;
  const int ke = min(kb + kBlockNnz, ptr[row + 1]);
  double s = 0.0;
  for (int k = kb + (int)threadIdx.x; k < ke; k += kThreads) s += (double)val[k] * (double)xv[col[k]];
  return block_sum(s, scratch);
}

// Thread 0 of a chunk block publishes its partial; returns true in the last
// chunk to finish, with *total the chunk sums combined in chunk order.
__device__ __forceinline__ bool longrow_publish(const LongRow lr, int chunk, int lid,
                                                int32_t *long_cnt, double *long_part,
                                                double s, double *total) {
  __hip_atomic_store(&long_part[lr.slot + chunk], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int prev = __hip_atomic_fetch_add(&long_cnt[lid], 1, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT);
  if (prev != lr.nchunks - 1) return false;
  double t = 0.0;
  for (int c = 0; c < lr.nchunks; ++c)
    t += __hip_atomic_load(&long_part[lr.slot + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&long_cnt[lid], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  *total = t;
  return true;
}

// ---------------------------------------------------------------------------
// Primal half-step over the rows of A^T.
// FLUSH: only the pending weighted sums and A^T y (evaluation preparation).
template <typename T, bool FLUSH>
__global__ __launch_bounds__(kThreads) void k_primal(const BlockDesc *__restrict__ blocks,
                                                     const LongRow *__restrict__ longs,
                                                     int32_t *__restrict__ long_cnt,
                                                     double *__restrict__ long_part,
                                                     const InstDev *__restrict__ insts,
                                                     const Ctl *__restrict__ ctl,
                                                     double *__restrict__ part) {
  __shared__ double lds[kBlockNnz];
  __shared__ double scratch[kThreads / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  const Ctl *C = &ctl[bd.inst];
  const bool run = FLUSH ? (C->eval_flag != 0) : (C->active != 0);
  if (!run) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  const bool pending = C->pending != 0;
  const bool need = C->aty_valid == 0;
  const double w = C->w_pending;
  const double tau = C->eta / C->omega;
  const T *__restrict__ yv = (const T *)(cur ? I.y1 : I.y0);
  const T *__restrict__ xv = (const T *)(cur ? I.x1 : I.x0);
  T *__restrict__ xn = (T *)(cur ? I.x0 : I.x1);
  T *__restrict__ aty = (T *)I.aty;
  T *__restrict__ dx = (T *)I.dx;
  T *__restrict__ sx = (T *)I.sx;
  T *__restrict__ saty = (T *)I.saty;
  const T *__restrict__ cv = (const T *)I.c;
  const T *__restrict__ lv = (const T *)I.l;
  const T *__restrict__ uv = (const T *)I.u;
  double acc = 0.0;
  auto epi = [&](int j, double s) {
    const double a = need ? s : (double)aty[j];
    const double xj = (double)xv[j];
    if (pending) {
      sx[j] = (T)((double)sx[j] + xj * w);
      saty[j] = (T)((double)saty[j] + a * w);
    }
    if (need) aty[j] = (T)a;
    if (!FLUSH) {
      double t = xj - tau * ((double)cv[j] - a);
      t = jl_min((double)uv[j], jl_max((double)lv[j], t));
      const double d = t - xj;
      dx[j] = (T)d;
      xn[j] = (T)(xj + d);
      acc += d * d;
    }
  };
  if (bd.nrows > 0) {
    rowblock_spmv(I.t_ptr, I.t_col, (const T *)I.t_val, yv, bd.row0, bd.nrows, lds, need, epi);
  } else {
    const int chunk = -bd.nrows - 1;
    if (!need) {
      if (chunk == 0 && threadIdx.x == 0) epi(bd.row0, 0.0);
    } else {
      const double s = longrow_chunk(I.t_ptr, I.t_col, (const T *)I.t_val, yv, bd.row0, chunk,
                                     scratch);
      if (threadIdx.x == 0) {
        double tot;
        if (longrow_publish(longs[bd.aux], chunk, bd.aux, long_cnt, long_part, s, &tot))
          epi(bd.row0, tot);
      }
    }
  }
  if (!FLUSH) {
    const double t = block_sum(acc, scratch);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// Dual half-step over the rows of A.  FLUSH: only the pending weighted sums.
template <typename T, bool FLUSH>
__global__ __launch_bounds__(kThreads) void k_dual(const BlockDesc *__restrict__ blocks,
                                                   const LongRow *__restrict__ longs,
                                                   int32_t *__restrict__ long_cnt,
                                                   double *__restrict__ long_part,
                                                   const InstDev *__restrict__ insts,
                                                   const Ctl *__restrict__ ctl,
                                                   double *__restrict__ part) {
  __shared__ double lds[kBlockNnz];
  __shared__ double scratch[kThreads / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  const Ctl *C = &ctl[bd.inst];
  const bool run = FLUSH ? (C->eval_flag != 0) : (C->active != 0);
  if (!run) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  const bool pending = C->pending != 0;
  const double w = C->w_pending;
  const double sig = C->omega * C->eta;
  const int neq = I.neq;
  const T *__restrict__ yv = (const T *)(cur ? I.y1 : I.y0);
  T *__restrict__ yn = (T *)(cur ? I.y0 : I.y1);
  const T *__restrict__ axv = (const T *)(cur ? I.ax1 : I.ax0);
  T *__restrict__ axn = (T *)(cur ? I.ax0 : I.ax1);
  T *__restrict__ sy = (T *)I.sy;
  T *__restrict__ sax = (T *)I.sax;
  const T *__restrict__ bv = (const T *)I.b;
  const T *__restrict__ dxv = (const T *)I.dx;
  double acc_dy = 0.0, acc_int = 0.0;
  auto epi = [&](int i, double adx) {
    const double yi = (double)yv[i];
    const double axi = (double)axv[i];
    if (pending) {
      sy[i] = (T)((double)sy[i] + yi * w);
      sax[i] = (T)((double)sax[i] + axi * w);
    }
    if (!FLUSH) {
      double t = yi + sig * (((double)bv[i] - 2.0 * adx) - axi);
      if (i >= neq) t = jl_max(t, 0.0);
      const double d = t - yi;
      yn[i] = (T)(yi + d);
      axn[i] = (T)(axi + adx);
      acc_dy += d * d;
      acc_int += adx * d;
    }
  };
  if (bd.nrows > 0) {
    rowblock_spmv(I.a_ptr, I.a_col, (const T *)I.a_val, dxv, bd.row0, bd.nrows, lds, !FLUSH, epi);
  } else {
    const int chunk = -bd.nrows - 1;
    if (FLUSH) {
      if (chunk == 0 && threadIdx.x == 0) epi(bd.row0, 0.0);
    } else {
      const double s = longrow_chunk(I.a_ptr, I.a_col, (const T *)I.a_val, dxv, bd.row0, chunk,
                                     scratch);
      if (threadIdx.x == 0) {
        double tot;
        if (longrow_publish(longs[bd.aux], chunk, bd.aux, long_cnt, long_part, s, &tot))
          epi(bd.row0, tot);
      }
    }
  }
  if (!FLUSH) {
    const double a = block_sum(acc_dy, scratch);
    const double b = block_sum(acc_int, scratch);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = a;
      part[2 * blockIdx.x + 1] = b;
    }
  }
}

// Reduction of the step partials and the adaptive step-size rule, one
// work-group per instance (src/primal_dual_hybrid_gradient_gpu.jl:324-412).
__global__ __launch_bounds__(kThreads) void k_step(const InstDev *__restrict__ insts,
                                                   Ctl *__restrict__ ctl,
                                                   const double *__restrict__ ppart,
                                                   const double *__restrict__ dpart,
                                                   const double *__restrict__ red_tab,
                                                   const double *__restrict__ grow_tab,
                                                   int64_t tab_len, double red_exp,
                                                   double grow_exp) {
  __shared__ double scratch[kThreads / 64];
  const int q = blockIdx.x;
  Ctl *C = &ctl[q];
  const int active = C->active;
  if (!active) return;
  const InstDev &I = insts[q];
  double sx = 0.0, sy = 0.0, sxy = 0.0;
  for (int b = I.pb0 + (int)threadIdx.x; b < I.pb1; b += kThreads) sx += ppart[b];
  for (int b = I.db0 + (int)threadIdx.x; b < I.db1; b += kThreads) {
    sy += dpart[2 * b];
    sxy += dpart[2 * b + 1];
  }
  sx = block_sum(sx, scratch);
  sy = block_sum(sy, scratch);
  sxy = block_sum(sxy, scratch);
  if (threadIdx.x != 0) return;
  const double omega = C->omega;
  double eta = C->eta;
  const int64_t k = C->attempts + 1;
  C->attempts = k;
  C->kkt_passes += 1.0;
  const double interaction = fabs(sxy);
  const double nx = sqrt(sx), ny = sqrt(sy);
  const double movement = 0.5 * omega * (nx * nx) + (0.5 / omega) * (ny * ny);
  C->last_interaction = interaction;
  C->last_movement = movement;
  double limit;
  if (interaction > 0.0) {
    limit = movement / interaction;
    if (movement == 0.0) {
      C->numerical_error = 1;
      C->active = 0;
      C->eta_entry = eta;
      return;
    }
  } else {
    limit = INFINITY;
  }
  const bool accept = eta <= limit;
  if (accept) {
    C->cur ^= 1;
    C->pending = 1;
    C->w_pending = C->eta_entry;
    C->wsum += C->eta_entry;
    C->count += 1;
    C->accepted += 1;
    C->aty_valid = 0;
  }
  double fr, fg;
  if (k < tab_len) {
    fr = red_tab[k];
    fg = grow_tab[k];
  } else {
    const double k1 = (double)(k + 1);
    fr = 1.0 - 1.0 / pow(k1, red_exp);
    fg = 1.0 + 1.0 / pow(k1, grow_exp);
  }
  eta = jl_min(fr * limit, fg * eta);
  C->eta = eta;
  if (accept) {
    C->eta_entry = eta;
    if (C->accepted >= C->target) C->active = 0;
  }
}

// ---------------------------------------------------------------------------
// Evaluation: averages, unscaled KKT statistics of the average
// (src/iteration_stats_utils_gpu.jl:26-502), the three weighted-KKT
// candidates of the restart scheme (src/saddle_point_gpu.jl:250-278, 479-622)
// and the restart distances (:673-682), all in one pass.  Blocks [0, np) walk
// the A^T row blocks (variables), the rest walk the A row blocks (rows).

__device__ __forceinline__ bool eval_is_max_n(int k) {
  return k == 1 || k == 3 || k == 6 || k == 8 || k == 10 || k == 11;
}
__device__ __forceinline__ bool eval_is_max_m(int k) {
  return k == 1 || k == 3 || k == 5 || k == 7;
}

__device__ __forceinline__ double contrib(double l, double u, double r) {
  return r > 0.0 ? l * r : (r < 0.0 ? u * r : 0.0);
}

template <int K, bool M>
__device__ __forceinline__ void block_reduce_acc(double (&acc)[K], double *lds, double *out) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool ismax = M ? eval_is_max_m(k) : eval_is_max_n(k);
    const double v = ismax ? wave_max(acc[k]) : wave_sum(acc[k]);
    if ((threadIdx.x & 63) == 0) lds[k * (kThreads / 64) + w] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const bool ismax = M ? eval_is_max_m(k) : eval_is_max_n(k);
    double s = lds[k * (kThreads / 64)];
    for (int q = 1; q < kThreads / 64; ++q)
      s = ismax ? jl_max(s, lds[k * (kThreads / 64) + q]) : s + lds[k * (kThreads / 64) + q];
    out[k] = s;
  }
}

__device__ __forceinline__ void block_rows(const BlockDesc &bd, int *r0, int *r1) {
  if (bd.nrows > 0) {
    *r0 = bd.row0;
    *r1 = bd.row0 + bd.nrows;
  } else if (bd.nrows == -1) {
    *r0 = bd.row0;
    *r1 = bd.row0 + 1;
  } else {
    *r0 = *r1 = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_eval(const BlockDesc *__restrict__ pblocks, int np,
                                                   const BlockDesc *__restrict__ dblocks,
                                                   const InstDev *__restrict__ insts,
                                                   const Ctl *__restrict__ ctl,
                                                   double *__restrict__ epn,
                                                   double *__restrict__ epm) {
  __shared__ double lds[kEvalN * (kThreads / 64)];
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  const Ctl *C = &ctl[bd.inst];
  if (!C->eval_flag) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  const bool usecur = C->use_current != 0;
  const double wsum = C->wsum;
  int r0, r1;
  block_rows(bd, &r0, &r1);
  if (nside) {
    double acc[kEvalN];
#pragma unroll
    for (int k = 0; k < kEvalN; ++k) acc[k] = 0.0;
    const void *xcur = cur ? I.x1 : I.x0;
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      const double xc = ld<T>(xcur, j);
      const double cs = ld<T>(I.c, j), ls = ld<T>(I.l, j), us = ld<T>(I.u, j);
      const double ax = usecur ? xc : ld<T>(I.sx, j) / wsum;
      const double ag = usecur ? ld<T>(I.gbuf, j) : (-ld<T>(I.saty, j)) / wsum + cs;
      const double gnew = cs - ld<T>(I.aty, j);
      st<T>(I.avg_x, j, ax);
      st<T>(I.avg_g, j, ag);
      st<T>(I.gbuf, j, gnew);
      // original space
      const double Dj = I.D[j];
      const double ox = ax / Dj, og = ag * Dj;
      const double ol = I.ol[j], ou = I.ou[j], oc = I.oc[j];
      const double fl = isfinite(ol) ? 1.0 : 0.0, fu = isfinite(ou) ? 1.0 : 0.0;
      const double lvv = jl_max(ol - ox, 0.0), uvv = jl_max(ox - ou, 0.0);
      acc[0] += lvv * lvv;
      acc[0] += uvv * uvv;
      acc[1] = jl_max(acc[1], jl_max(lvv, uvv));
      acc[2] += oc * ox;
      acc[3] = jl_max(acc[3], fabs(ox));
      acc[4] += ox * ox;
      const double r = jl_max(og, 0.0) * fl + jl_min(og, 0.0) * fu;
      const double rv = og - r;
      acc[5] += rv * rv;
      acc[6] = jl_max(acc[6], fabs(rv));
      acc[7] += contrib(ol, ou, r);
      const double g2 = og - oc;
      const double r2 = jl_max(g2, 0.0) * fl + jl_min(g2, 0.0) * fu;
      const double rv2 = g2 - r2;
      acc[8] = jl_max(acc[8], fabs(rv2));
      acc[9] += contrib(ol, ou, r2);
      acc[10] = jl_max(acc[10], fabs(r2));
      acc[11] = jl_max(acc[11], jl_max(fl != 0.0 ? jl_max(-ox, 0.0) : 0.0,
                                        fu != 0.0 ? jl_max(ox, 0.0) : 0.0));
      // restart candidates in the scaled space: current, average, last restart
      const double xs3[3] = {xc, ax, ld<T>(I.lx, j)};
      const double gs3[3] = {gnew, ag, ld<T>(I.lg, j)};
      const double sfl = isfinite(ls) ? 1.0 : 0.0, sfu = isfinite(us) ? 1.0 : 0.0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double xs = xs3[q], gs = gs3[q];
        const double a = jl_max(ls - xs, 0.0), b = jl_max(xs - us, 0.0);
        acc[12 + 4 * q] += a * a;
        acc[12 + 4 * q] += b * b;
        acc[13 + 4 * q] += cs * xs;
        const double rr = jl_max(gs, 0.0) * sfl + jl_min(gs, 0.0) * sfu;
        const double rvv = gs - rr;
        acc[14 + 4 * q] += rvv * rvv;
        acc[15 + 4 * q] += contrib(ls, us, rr);
      }
      const double dxl = ax - xs3[2];
      acc[24] += dxl * dxl;
    }
    block_reduce_acc<kEvalN, false>(acc, lds, epn + (size_t)blk * kEvalN);
  } else {
    double acc[kEvalM];
#pragma unroll
    for (int k = 0; k < kEvalM; ++k) acc[k] = 0.0;
    const void *ycur = cur ? I.y1 : I.y0;
    const void *axcur = cur ? I.ax1 : I.ax0;
    const int neq = I.neq;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      const double yc = ld<T>(ycur, i), axc = ld<T>(axcur, i);
      const double ay = usecur ? yc : ld<T>(I.sy, i) / wsum;
      const double aax = usecur ? axc : ld<T>(I.sax, i) / wsum;
      st<T>(I.avg_y, i, ay);
      st<T>(I.avg_ax, i, aax);
      const double Ei = I.E[i];
      const double oy = ay / Ei, oax = aax * Ei;
      const double ob = I.ob[i];
      const double t = ob - oax;
      const double cvv = i < neq ? t : jl_max(t, 0.0);
      acc[0] += cvv * cvv;
      acc[1] = jl_max(acc[1], fabs(cvv));
      if (i >= neq) {
        const double dr = jl_max(-oy, 0.0);
        acc[2] += dr * dr;
        acc[3] = jl_max(acc[3], dr);
      }
      acc[4] += ob * oy;
      acc[5] = jl_max(acc[5], fabs(oy));
      acc[6] += oy * oy;
      acc[7] = jl_max(acc[7], i < neq ? fabs(oax) : jl_max(-oax, 0.0));
      const double bs = ld<T>(I.b, i);
      const double ys3[3] = {yc, ay, ld<T>(I.ly, i)};
      const double as3[3] = {axc, aax, ld<T>(I.lax, i)};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double tt = bs - as3[q];
        const double cc = i < neq ? tt : jl_max(tt, 0.0);
        acc[8 + 3 * q] += cc * cc;
        if (i >= neq) {
          const double dr = jl_max(-ys3[q], 0.0);
          acc[9 + 3 * q] += dr * dr;
        }
        acc[10 + 3 * q] += bs * ys3[q];
      }
      const double dyl = ay - ys3[2];
      acc[17] += dyl * dyl;
    }
    block_reduce_acc<kEvalM, true>(acc, lds, epm + (size_t)blk * kEvalM);
  }
}

// Per-instance combination of the evaluation partials: one work-group per
// (instance, accumulator), blocks combined in a fixed strided order.
__global__ __launch_bounds__(kThreads) void k_eval_finish(const InstDev *__restrict__ insts,
                                                          const Ctl *__restrict__ ctl,
                                                          const double *__restrict__ epn,
                                                          const double *__restrict__ epm,
                                                          double *__restrict__ out) {
  __shared__ double scratch[kThreads / 64];
  const int q = blockIdx.x;
  const int k = blockIdx.y;
  if (!ctl[q].eval_flag) return;
  const InstDev &I = insts[q];
  const bool nside = k < kEvalN;
  const int kk = nside ? k : k - kEvalN;
  const int K = nside ? kEvalN : kEvalM;
  const bool ismax = nside ? eval_is_max_n(kk) : eval_is_max_m(kk);
  const double *ep = nside ? epn : epm;
  const int b0 = nside ? I.pb0 : I.db0, b1 = nside ? I.pb1 : I.db1;
  double s = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += kThreads) {
    const double v = ep[(size_t)b * K + kk];
    s = ismax ? jl_max(s, v) : s + v;
  }
  s = ismax ? wave_max(s) : wave_sum(s);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = scratch[0];
    for (int i = 1; i < kThreads / 64; ++i) t = ismax ? jl_max(t, scratch[i]) : t + scratch[i];
    out[(size_t)q * kEvalOut + k] = t;
  }
}

// Restart actions (src/saddle_point_gpu.jl:572-622, 659-692): 1 = reset the
// average at the current point, 2 = restart to the average.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_restart(const BlockDesc *__restrict__ pblocks,
                                                      int np,
                                                      const BlockDesc *__restrict__ dblocks,
                                                      const InstDev *__restrict__ insts,
                                                      const Ctl *__restrict__ ctl) {
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  const Ctl *C = &ctl[bd.inst];
  const int action = C->restart_action;
  if (!C->eval_flag || action == 0) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  int r0, r1;
  block_rows(bd, &r0, &r1);
  if (nside) {
    T *xc = (T *)(cur ? I.x1 : I.x0);
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      if (action == 2) {
        const double ag = ld<T>(I.avg_g, j);
        xc[j] = ((const T *)I.avg_x)[j];
        st<T>(I.aty, j, ld<T>(I.c, j) - ag);
        st<T>(I.gbuf, j, ag);
      }
      ((T *)I.sx)[j] = (T)0;
      ((T *)I.saty)[j] = (T)0;
      ((T *)I.lx)[j] = xc[j];
      ((T *)I.lg)[j] = ((const T *)I.gbuf)[j];
    }
  } else {
    T *yc = (T *)(cur ? I.y1 : I.y0);
    T *axc = (T *)(cur ? I.ax1 : I.ax0);
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      if (action == 2) {
        yc[i] = ((const T *)I.avg_y)[i];
        axc[i] = ((const T *)I.avg_ax)[i];
      }
      ((T *)I.sy)[i] = (T)0;
      ((T *)I.sax)[i] = (T)0;
      ((T *)I.ly)[i] = yc[i];
      ((T *)I.lax)[i] = axc[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers.

static inline unsigned grid_of(int n) { return (unsigned)(n > 0 ? n : 0); }

#define PDLP_CHECK_LAUNCH() return (int)hipGetLastError()

int launch_primal(hipStream_t s, int prec, bool flush, int nblocks, const BlockDesc *blocks,
                  const LongRow *longs, int32_t *long_cnt, double *long_part,
                  const InstDev *insts, const Ctl *ctl, double *part) {
  if (nblocks <= 0) return 0;
  dim3 g(grid_of(nblocks)), b(kThreads);
  if (prec == 0) {
    if (flush)
      hipLaunchKernelGGL((k_primal<double, true>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
    else
      hipLaunchKernelGGL((k_primal<double, false>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
  } else {
    if (flush)
      hipLaunchKernelGGL((k_primal<float, true>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
    else
      hipLaunchKernelGGL((k_primal<float, false>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
  }
  PDLP_CHECK_LAUNCH();
}

int launch_dual(hipStream_t s, int prec, bool flush, int nblocks, const BlockDesc *blocks,
                const LongRow *longs, int32_t *long_cnt, double *long_part, const InstDev *insts,
                const Ctl *ctl, double *part) {
  if (nblocks <= 0) return 0;
  dim3 g(grid_of(nblocks)), b(kThreads);
  if (prec == 0) {
    if (flush)
      hipLaunchKernelGGL((k_dual<double, true>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
    else
      hipLaunchKernelGGL((k_dual<double, false>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
  } else {
    if (flush)
      hipLaunchKernelGGL((k_dual<float, true>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
    else
      hipLaunchKernelGGL((k_dual<float, false>), g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl, part);
  }
  PDLP_CHECK_LAUNCH();
}

int launch_step(hipStream_t s, int ninst, const InstDev *insts, Ctl *ctl, const double *ppart,
                const double *dpart, const double *red_tab, const double *grow_tab,
                int64_t tab_len, double red_exp, double grow_exp) {
  if (ninst <= 0) return 0;
  hipLaunchKernelGGL(k_step, dim3(grid_of(ninst)), dim3(kThreads), 0, s, insts, ctl, ppart, dpart,
                     red_tab, grow_tab, tab_len, red_exp, grow_exp);
  PDLP_CHECK_LAUNCH();
}

int launch_eval(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl, double *epn,
                double *epm, int ninst, double *out) {
  if (np + nd > 0) {
    if (prec == 0)
      hipLaunchKernelGGL(k_eval<double>, dim3(grid_of(np + nd)), dim3(kThreads), 0, s, pblocks, np,
                         dblocks, insts, ctl, epn, epm);
    else
      hipLaunchKernelGGL(k_eval<float>, dim3(grid_of(np + nd)), dim3(kThreads), 0, s, pblocks, np,
                         dblocks, insts, ctl, epn, epm);
  }
  if (ninst > 0)
    hipLaunchKernelGGL(k_eval_finish, dim3(grid_of(ninst), kEvalOut), dim3(kThreads), 0, s, insts,
                       ctl, epn, epm, out);
  PDLP_CHECK_LAUNCH();
}

int launch_restart(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                   const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl) {
  if (np + nd <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_restart<double>, dim3(grid_of(np + nd)), dim3(kThreads), 0, s, pblocks,
                       np, dblocks, insts, ctl);
  else
    hipLaunchKernelGGL(k_restart<float>, dim3(grid_of(np + nd)), dim3(kThreads), 0, s, pblocks,
                       np, dblocks, insts, ctl);
  PDLP_CHECK_LAUNCH();
}

}  // namespace pdlp
