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
// Every sum (SpMV rows and all reductions) in the FP64 path is a compensated
// sum of error-free products (TwoProd with FMA, TwoSum) whose rounding errors
// are themselves added in plain FP64 (the Dot2 scheme of Ogita, Rump and
// Oishi): |result - exact| <= eps |exact| + gamma_n^2 sum |terms|.  Unless a
// sum cancels to ~n^2 eps of its terms' magnitude, that is the FP64 value
// nearest the exact sum whatever the summation order; under heavier
// cancellation two orders can differ in the last bits.  Bitwise agreement
// with the CPU oracle (which sums in another order) is therefore what the
// tests observe, not a theorem.  Sums are deterministic: per-thread sums in a
// fixed stride order, wave butterflies, waves and blocks combined in order,
// and an instance's sums do not depend on which other instances share the
// launch.
//
// Floating-point contraction is disabled for this file (-ffp-contract=off) so
// each elementwise formula rounds exactly as the reference's written form.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

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

// Compensated accumulator: value s + e, with e collecting the rounding errors.
struct DD {
  double s, e;
};
__device__ __forceinline__ DD dd_zero() { return DD{0.0, 0.0}; }
// s += v, error into e (TwoSum)
__device__ __forceinline__ void dd_add(DD &a, double v) {
  const double s = a.s + v;
  const double bb = s - a.s;
  const double t = (a.s - (s - bb)) + (v - bb);
  a.s = s;
  a.e += t;
}
// s += x*y exactly (TwoProd via FMA, then TwoSum)
__device__ __forceinline__ void dd_add_prod(DD &a, double x, double y) {
  const double p = x * y;
  const double q = __builtin_fma(x, y, -p);
  dd_add(a, p);
  a.e += q;
}
__device__ __forceinline__ void dd_merge(DD &a, const DD &b) {
  const double s = a.s + b.s;
  const double bb = s - a.s;
  const double t = (a.s - (s - bb)) + (b.s - bb);
  a.s = s;
  a.e = (a.e + b.e) + t;
}
__device__ __forceinline__ double dd_val(const DD &a) { return a.s + a.e; }

__device__ __forceinline__ DD wave_dd(DD v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    DD w{__shfl_xor(v.s, o, 64), __shfl_xor(v.e, o, 64)};
    dd_merge(v, w);
  }
  return v;
}

// Compensated sum over the work-group; valid in thread 0.  `scratch` holds
// 2*kThreads/64 doubles.  Contains barriers.
__device__ __forceinline__ DD block_dd(DD v, double *scratch) {
  v = wave_dd(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    scratch[2 * w] = v.s;
    scratch[2 * w + 1] = v.e;
  }
  __syncthreads();
  DD r = dd_zero();
  if (threadIdx.x == 0) {
    r = DD{scratch[0], scratch[1]};
#pragma unroll
    for (int k = 1; k < kThreads / 64; ++k) dd_merge(r, DD{scratch[2 * k], scratch[2 * k + 1]});
  }
  return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = jl_max(v, __shfl_xor(v, o, 64));
  return v;
}

// Global-address-space views of device pointers.  Pointers read from the
// instance tables are generic, and loads through them compile to flat
// loads, which also count against lgkmcnt: every LDS or cross-lane wait then
// waits for the flat loads in flight as well.
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;
template <typename T>
using gptr_w = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gp(const T *p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr_w<T> gpw(T *p) {
  return (gptr_w<T>)p;
}

template <typename T>
__device__ __forceinline__ double ld(const void *p, int i) {
  return (double)gp((const T *)p)[i];
}
template <typename T>
__device__ __forceinline__ void st(void *p, int i, double v) {
  gpw((T *)p)[i] = (T)v;
}

// Row sums of one row block.  Phase 1: the block's nonzeros are read as one
// flat, fully coalesced range (non-temporal loads for the streamed matrix, so
// the gathered vector keeps the caches), each exact product (hi, lo) is
// staged in LDS.  Phase 2: groups of `tpr` lanes (a power of two, at most a
// wave) form each row's compensated sum into rowsum[r].  The caller then runs
// the elementwise epilogue with consecutive threads on consecutive rows.
template <typename T, typename V>
__device__ __forceinline__ void rowblock_sums(const int32_t *__restrict__ ptr,
                                              const int32_t *__restrict__ col,
                                              const T *__restrict__ val,
                                              const V *__restrict__ xv, int row0, int nrows,
                                              double *lds_p, double *lds_q, double *rowsum) {
  const gptr<int32_t> pg = gp(ptr), cg = gp(col);
  const gptr<T> vg = gp(val);
  const gptr<V> xg = gp(xv);
  const int k0 = pg[row0];
  const int nz = pg[row0 + nrows] - k0;
#pragma unroll 4
  for (int k = threadIdx.x; k < nz; k += kThreads) {
    const double a = (double)__builtin_nontemporal_load(vg + k0 + k);
    const double x = (double)xg[__builtin_nontemporal_load(cg + k0 + k)];
    const double pr = a * x;
    lds_p[k] = pr;
    lds_q[k] = __builtin_fma(a, x, -pr);
  }
  __syncthreads();
  int tpr = 1;
  while (tpr < 64 && tpr * 2 * nrows <= kThreads) tpr <<= 1;
  const int lane = threadIdx.x & (tpr - 1);
  const int grp = threadIdx.x / tpr;
  const int ngrp = kThreads / tpr;
  for (int r = grp; r < nrows; r += ngrp) {
    const int rs = pg[row0 + r] - k0, re = pg[row0 + r + 1] - k0;
    DD acc = dd_zero();
    for (int k = rs + lane; k < re; k += tpr) {
      dd_add(acc, lds_p[k]);
      acc.e += lds_q[k];
    }
    for (int o = tpr >> 1; o > 0; o >>= 1) {
      const DD w{__shfl_xor(acc.s, o, tpr), __shfl_xor(acc.e, o, tpr)};
      dd_merge(acc, w);
    }
    if (lane == 0) rowsum[r] = dd_val(acc);
  }
  __syncthreads();
}

// Compensated partial sum of one kBlockNnz chunk of a long row; valid in thread 0.
template <typename T, typename V>
__device__ __forceinline__ DD longrow_chunk(const int32_t *__restrict__ ptr,
                                            const int32_t *__restrict__ col,
                                            const T *__restrict__ val,
                                            const V *__restrict__ xv, int row, int chunk,
                                            double *scratch) {
  const int kb = ptr[row] + chunk * kLongChunk;
  const int ke = min(kb + kLongChunk, ptr[row + 1]);
  DD acc = dd_zero();
  for (int k = kb + (int)threadIdx.x; k < ke; k += kThreads)
    dd_add_prod(acc, (double)__builtin_nontemporal_load(&val[k]),
                (double)xv[__builtin_nontemporal_load(&col[k])]);
  return block_dd(acc, scratch);
}

// Thread 0 of a chunk block publishes its partial; returns true in the last
// chunk to finish, with *total the chunk sums combined in chunk order.
__device__ __forceinline__ bool longrow_publish(const LongRow lr, int chunk, int lid,
                                                int32_t *long_cnt, double *long_part, DD s,
                                                double *total) {
  if (lr.nchunks == 1) {  // whole row in this work-group: no handoff
    *total = dd_val(s);
    return true;
  }
  __hip_atomic_store(&long_part[2 * (lr.slot + chunk)], s.s, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&long_part[2 * (lr.slot + chunk) + 1], s.e, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  const int prev = __hip_atomic_fetch_add(&long_cnt[lid], 1, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT);
  if (prev != lr.nchunks - 1) return false;
  DD t = dd_zero();
  for (int c = 0; c < lr.nchunks; ++c) {
    const DD v{__hip_atomic_load(&long_part[2 * (lr.slot + c)], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT),
               __hip_atomic_load(&long_part[2 * (lr.slot + c) + 1], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT)};
    dd_merge(t, v);
  }
  __hip_atomic_store(&long_cnt[lid], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  *total = dd_val(t);
  return true;
}

// ---------------------------------------------------------------------------
// Evaluation preparation on the A^T row blocks: adds the pending weighted
// sums of x and A^T y and forms A^T y of the current y when it is not valid.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_primal_flush(const BlockDesc *__restrict__ blocks,
                                                           const LongRow *__restrict__ longs,
                                                           int32_t *__restrict__ long_cnt,
                                                           double *__restrict__ long_part,
                                                           const InstDev *__restrict__ insts,
                                                           const Ctl *__restrict__ ctl) {
  __shared__ double lds_p[kBlockNnz];
  __shared__ double lds_q[kBlockNnz];
  __shared__ double rowsum[kBlockRows];
  __shared__ double scratch[2 * kThreads / 64];
  const BlockDesc bd = blocks[blockIdx.x];
  const Ctl *C = &ctl[bd.inst];
  if (C->eval_flag == 0) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  const bool pending = C->pending != 0;
  const bool need = C->aty_valid == 0;
  const double w = C->w_pending;
  const T *__restrict__ yv = (const T *)(cur ? I.y1 : I.y0);
  const T *__restrict__ xv = (const T *)(cur ? I.x1 : I.x0);
  T *__restrict__ aty = (T *)I.aty;
  double *__restrict__ sx = (double *)I.sx;
  double *__restrict__ saty = (double *)I.saty;
  auto epi = [&](int j, double s) {
    const double a = need ? s : (double)aty[j];
    if (pending) {
      sx[j] = sx[j] + (double)xv[j] * w;
      saty[j] = saty[j] + a * w;
    }
    if (need) aty[j] = (T)a;
  };
  if (bd.nrows >= 0) {
    if (need)
      rowblock_sums(I.t_ptr, I.t_col, (const T *)I.t_val, yv, bd.row0, bd.nrows, lds_p, lds_q,
                    rowsum);
    for (int r = threadIdx.x; r < bd.nrows; r += kThreads) epi(bd.row0 + r, need ? rowsum[r] : 0.0);
  } else {
    const int chunk = -bd.nrows - 1;
    if (!need) {
      if (chunk == 0 && threadIdx.x == 0) epi(bd.row0, 0.0);
    } else {
      const DD sm = longrow_chunk(I.t_ptr, I.t_col, (const T *)I.t_val, yv, bd.row0, chunk, scratch);
      if (threadIdx.x == 0) {
        double tot;
        if (longrow_publish(longs[bd.aux], chunk, bd.aux, long_cnt, long_part, sm, &tot))
          epi(bd.row0, tot);
      }
    }
  }
}

// One lane's share of a slice row: entries off + lane + 64 k, k < width,
// gathered from xv, as a compensated sum.  Padding slots (column -1) are
// skipped.  Entries go in groups of U whose index and value loads are all
// issued before their gathers; the last group is predicated.  U changes only
// how many loads are in flight, not the order of the sum.
// s += x*y for x, y widened from FP32: the product of two floats is exact in
// FP64 (48 significant bits, and the exponent range of a double holds every
// product of two floats), so TwoProd's error term is +0 and is not formed.
template <typename T, typename V>
__device__ __forceinline__ void dd_add_prod_of(DD &a, T x, V y) {
  if constexpr (std::is_same<T, float>::value && std::is_same<V, float>::value)
    dd_add(a, (double)x * (double)y);
  else
    dd_add_prod(a, (double)x, (double)y);
}

template <typename T, typename V, int U = 8>
__device__ __forceinline__ DD slice_lane_sum(const int32_t *__restrict__ scol,
                                             const T *__restrict__ sval,
                                             gptr<V> xg, int off, int width, int lane) {
  DD acc = dd_zero();
  const gptr<int32_t> cp = gp(scol + off + lane);
  const gptr<T> vp = gp(sval + off + lane);
  int k = 0;
  for (; k + U <= width; k += U) {
    int32_t c[U];
    T v[U];
    V g[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      c[i] = __builtin_nontemporal_load(cp + 64 * (k + i));
      v[i] = __builtin_nontemporal_load(vp + 64 * (k + i));
    }
#pragma unroll
    for (int i = 0; i < U; ++i) g[i] = c[i] >= 0 ? xg[c[i]] : V(0);
#pragma unroll
    for (int i = 0; i < U; ++i)
      if (c[i] >= 0) dd_add_prod_of(acc, v[i], g[i]);
  }
  if (k < width) {
    int32_t c[U];
    T v[U];
    V g[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const bool in = k + i < width;
      c[i] = in ? __builtin_nontemporal_load(cp + 64 * (k + i)) : -1;
      v[i] = in ? __builtin_nontemporal_load(vp + 64 * (k + i)) : T(0);
    }
#pragma unroll
    for (int i = 0; i < U; ++i) g[i] = c[i] >= 0 ? xg[c[i]] : V(0);
#pragma unroll
    for (int i = 0; i < U; ++i)
      if (c[i] >= 0) dd_add_prod_of(acc, v[i], g[i]);
  }
  return acc;
}

// Compensated sum over each group of 2^lg lanes (every lane of the group
// holds the result), rounded once.
__device__ __forceinline__ double group_sum(DD acc, int lg) {
  for (int o = (1 << lg) >> 1; o > 0; o >>= 1) {
    const DD w{__shfl_xor(acc.s, o, 64), __shfl_xor(acc.e, o, 64)};
    dd_merge(acc, w);
  }
  return dd_val(acc);
}

// The descriptor of wave-uniform slice wv as eight 32-bit words, loaded
// together (one scalar load), its 16-bit fields unpacked from the fifth
// (little-endian layout of SliceDesc: lg, then nrows).
__device__ __forceinline__ SliceDesc load_slice(const SliceDesc *slices, int wv) {
  static_assert(sizeof(SliceDesc) == 32, "SliceDesc is eight 32-bit words");
  const gptr<int32_t> w = (gptr<int32_t>)(slices + wv);
  const int32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
  const int32_t w4 = w[4], w5 = w[5], w6 = w[6], w7 = w[7];
  // no instruction: keeps the words from being loaded one test at a time
  asm volatile("" ::"s"(w0), "s"(w1), "s"(w2), "s"(w3), "s"(w4), "s"(w5), "s"(w6), "s"(w7));
  SliceDesc sd;
  sd.inst = w0;
  sd.row0 = w1;
  sd.off = w2;
  sd.width = w3;
  sd.lg = (int16_t)(w4 & 0xffff);
  sd.nrows = (int16_t)((uint32_t)w4 >> 16);
  sd.lid = w5;
  sd.chunk = w6;
  sd.part = w7;
  return sd;
}

// Per-launch view of one instance for the primal step.
template <typename T>
struct PrimalView {
  gptr<T> yv, xv, cv, lv, uv;
  gptr_w<T> xn, aty, dx;
  gptr_w<double> sx, saty;
  bool pending, need;
  double w, tau;
  __device__ PrimalView(const InstDev &I, const Ctl *C) {
    const int cur = C->cur;
    pending = C->pending != 0;
    need = C->aty_valid == 0;
    w = C->w_pending;
    tau = C->eta / C->omega;
    yv = gp((const T *)(cur ? I.y1 : I.y0));
    xv = gp((const T *)(cur ? I.x1 : I.x0));
    xn = gpw((T *)(cur ? I.x0 : I.x1));
    aty = gpw((T *)I.aty);
    dx = gpw((T *)I.dx);
    sx = gpw((double *)I.sx);
    saty = gpw((double *)I.saty);
    cv = gp((const T *)I.c);
    lv = gp((const T *)I.l);
    uv = gp((const T *)I.u);
  }
};

// Operands of one primal row.
struct PrimalRow {
  double x, c, l, u, sx, saty, aty;
};

template <typename T>
__device__ __forceinline__ PrimalRow primal_load(const PrimalView<T> &v, int j) {
  PrimalRow r;
  r.x = (double)v.xv[j];
  r.c = (double)v.cv[j];
  r.l = (double)v.lv[j];
  r.u = (double)v.uv[j];
  r.sx = v.pending ? v.sx[j] : 0.0;
  r.saty = v.pending ? v.saty[j] : 0.0;
  r.aty = v.need ? 0.0 : (double)v.aty[j];
  return r;
}

// The pending weighted-sum update and the projected primal step (K1,
// src/primal_dual_hybrid_gradient_gpu.jl:177-197) of row j, given the row
// sum s of A^T y (used when A^T y is recomputed); adds dx^2 to acc.
template <typename T>
__device__ __forceinline__ void primal_update(const PrimalView<T> &v, int j, const PrimalRow &r,
                                              double s, DD &acc) {
  const double a = v.need ? s : r.aty;
  if (v.pending) {
    v.sx[j] = r.sx + r.x * v.w;
    v.saty[j] = r.saty + a * v.w;
  }
  if (v.need) v.aty[j] = (T)a;
  double t = r.x - v.tau * (r.c - a);
  t = jl_min(r.u, jl_max(r.l, t));
  if constexpr (std::is_same<T, double>::value) {
    const double d = t - r.x;
    v.dx[j] = d;
    v.xn[j] = r.x + d;
    dd_add_prod(acc, d, d);
  } else {
    // FP32 iterates: the step is the change of the stored value (the dual
    // kernel multiplies A by the new point itself, not by dx)
    const T xr = (T)t;
    const double d = (double)xr - r.x;
    v.xn[j] = xr;
    dd_add_prod(acc, d, d);
  }
}

// Primal half-step over the slices of A^T: A^T y (only when y changed), the
// pending weighted-sum update of x and A^T y, the projected primal step and
// the partial sum of dx^2 of each slice (part[2p], part[2p+1] for the
// slice's partial index p).
// One wave per slice; each row's operands are loaded before the gathers of
// its sum.
template <typename T, int U = 8>
__device__ __forceinline__ void primal_slice(const SliceDesc &sd, int lane,
                                             const InstDev *__restrict__ insts, const Ctl *ctl,
                                             double *part) {
  if (sd.nrows < 0) return;
  const Ctl *C = &ctl[sd.inst];
  const int active = C->active;
  const InstDev &I = insts[sd.inst];
  const PrimalView<T> v(I, C);
  const int32_t *scol = I.ts_col;
  const T *sval = (const T *)I.ts_val;
  // issue the control block's and the instance's loads together (the empty
  // asm holds no instruction; it keeps them from being sunk below the test)
  asm volatile("" ::"s"(active), "s"(v.yv), "s"(v.xv), "s"(v.cv), "s"(v.lv), "s"(v.uv),
               "s"(v.xn), "s"(v.sx), "s"(v.saty), "s"(scol), "s"(sval), "s"(C->eta),
               "s"(C->omega), "s"(C->w_pending), "s"(C->pending), "s"(C->aty_valid));
  if (active == 0) return;
  const int j = sd.row0 + lane;
  const bool mine = lane < sd.nrows;
  PrimalRow r{};
  if (mine) r = primal_load(v, j);
  double s = 0.0;
  if (v.need) {
    s = group_sum(slice_lane_sum<T, T, U>(scol, sval, v.yv, sd.off, sd.width, lane), sd.lg);
    s = __shfl(s, (lane << sd.lg) & 63, 64);
  }
  DD acc = dd_zero();
  if (mine) primal_update(v, j, r, s, acc);
  acc = wave_dd(acc);
  if (lane == 0) {
    part[2 * sd.part] = acc.s;
    part[2 * sd.part + 1] = acc.e;
  }
}

// One wave per slice; W waves per work-group, U loads in flight per lane.
template <typename T, int U, int W>
__global__ __launch_bounds__(64 * W) void k_primal(const SliceDesc *__restrict__ slices, int nsl,
                                                   const InstDev *__restrict__ insts,
                                                   const Ctl *__restrict__ ctl,
                                                   double *__restrict__ part) {
  // wave-uniform: the descriptor is one scalar load, and the wave's
  // instance-wide values live in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane((int)blockIdx.x * W + ((int)threadIdx.x >> 6));
  if (wv >= nsl) return;
  primal_slice<T, U>(load_slice(slices, wv), threadIdx.x & 63, insts, ctl, part);
}

// The chunk slices of rows longer than kSellChunk entries: each wave sums one
// chunk; the last chunk of a row to finish combines the chunk sums in chunk
// order and updates the row.  part as in k_primal.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_primal_long(const SliceDesc *__restrict__ slices,
                                                          int nsl, const LongRow *__restrict__ longs,
                                                          int32_t *__restrict__ long_cnt,
                                                          double *__restrict__ long_part,
                                                          const InstDev *__restrict__ insts,
                                                          const Ctl *__restrict__ ctl,
                                                          double *__restrict__ part) {
  const int wv = (int)blockIdx.x * (kThreads / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = slices[wv];
  const Ctl *C = &ctl[sd.inst];
  if (C->active == 0) return;
  const InstDev &I = insts[sd.inst];
  const PrimalView<T> v(I, C);
  DD acc = dd_zero();
  if (v.need) {
    const DD t = wave_dd(slice_lane_sum(I.ts_col, (const T *)I.ts_val, v.yv, sd.off, sd.width, lane));
    if (lane == 0) {
      double tot;
      if (longrow_publish(longs[sd.lid], sd.chunk, sd.lid, long_cnt, long_part, t, &tot))
        primal_update(v, sd.row0, primal_load(v, sd.row0), tot, acc);
    }
  } else if (sd.chunk == 0 && lane == 0) {
    primal_update(v, sd.row0, primal_load(v, sd.row0), 0.0, acc);
  }
  if (lane == 0) {
    part[2 * sd.part] = acc.s;
    part[2 * sd.part + 1] = acc.e;
  }
}

// The adaptive step rule of take_step! (src/primal_dual_hybrid_gradient_gpu.jl:324-412)
// applied to the reduced sums of one attempt; runs on one thread.
// debug_attempt > 0: the attempt with that number sees movement = 0 (fault
// injection of the numerical-error branch for tests).
// fr_tab / fg_tab: the step factors of this attempt read from the host's
// tables when the attempt number is below tab_len.
__device__ __forceinline__ void step_rule(Ctl *C, double ssx, double ssy, double sxy,
                                          double fr_tab, double fg_tab, int64_t tab_len,
                                          double red_exp, double grow_exp,
                                          int64_t debug_attempt) {
  const double omega = C->omega;
  double eta = C->eta;
  const int64_t k = C->attempts + 1;
  C->attempts = k;
  C->kkt_passes += 1.0;
  // This attempt's passes added the pending weighted sums and left A^T y of
  // the current y in aty.
  C->pending = 0;
  C->aty_valid = 1;
  const double interaction = fabs(sxy);
  const double nx = sqrt(ssx), ny = sqrt(ssy);
  double movement = 0.5 * omega * (nx * nx) + (0.5 / omega) * (ny * ny);
  if (k == debug_attempt) movement = 0.0;
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
    fr = fr_tab;
    fg = fg_tab;
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

// Evaluation preparation on the A row blocks: adds the pending weighted sums
// of y and A x.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dual_flush(const BlockDesc *__restrict__ blocks,
                                                         const InstDev *__restrict__ insts,
                                                         const Ctl *__restrict__ ctl) {
  const BlockDesc bd = blocks[blockIdx.x];
  const Ctl *C = &ctl[bd.inst];
  if (C->eval_flag == 0 || C->pending == 0) return;
  const InstDev &I = insts[bd.inst];
  const int cur = C->cur;
  const double w = C->w_pending;
  const T *__restrict__ yv = (const T *)(cur ? I.y1 : I.y0);
  const double *__restrict__ axv = (const double *)(cur ? I.ax1 : I.ax0);
  double *__restrict__ sy = (double *)I.sy;
  double *__restrict__ sax = (double *)I.sax;
  int r0, r1;
  if (bd.nrows >= 0) {
    r0 = bd.row0;
    r1 = bd.row0 + bd.nrows;
  } else {
    if (bd.nrows != -1) return;  // chunk 0 of a long row handles it
    r0 = bd.row0;
    r1 = bd.row0 + 1;
  }
  for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
    sy[i] = sy[i] + (double)yv[i] * w;
    sax[i] = sax[i] + axv[i] * w;
  }
}

// Per-launch view of one instance for the dual step.  The dual kernel
// multiplies A by dx (FP64; the products are updated incrementally, as the
// reference does) or by the new point x' (FP32, so that A x never drifts
// from the stored point).
template <typename T>
struct DualView {
  gptr<T> yv, bv, dxv, ylo;
  gptr_w<T> yn;
  gptr<double> axv;
  gptr_w<double> axn, sy, sax;
  bool pending;
  double w, sig;
  int neq;
  __device__ DualView(const InstDev &I, const Ctl *C) {
    const int cur = C->cur;
    pending = C->pending != 0;
    w = C->w_pending;
    sig = C->omega * C->eta;
    neq = I.neq;
    yv = gp((const T *)(cur ? I.y1 : I.y0));
    yn = gpw((T *)(cur ? I.y0 : I.y1));
    axv = gp((const double *)(cur ? I.ax1 : I.ax0));
    axn = gpw((double *)(cur ? I.ax0 : I.ax1));
    sy = gpw((double *)I.sy);
    sax = gpw((double *)I.sax);
    bv = gp((const T *)I.b);
    dxv = gp((const T *)(std::is_same<T, float>::value ? (cur ? I.x0 : I.x1) : I.dx));
    ylo = gp((const T *)I.ylo);
  }
};

struct DualRow {
  double y, ax, b, sy, sax, ylo;
};

template <typename T>
__device__ __forceinline__ DualRow dual_load(const DualView<T> &v, int i) {
  DualRow r;
  r.y = (double)v.yv[i];
  r.ax = v.axv[i];
  r.b = (double)v.bv[i];
  r.sy = v.pending ? v.sy[i] : 0.0;
  r.sax = v.pending ? v.sax[i] : 0.0;
  r.ylo = (v.ylo && i >= v.neq) ? (double)v.ylo[i] : 0.0;
  return r;
}

// The pending weighted-sum update and the projected dual step with
// extrapolation (K2, src/primal_dual_hybrid_gradient_gpu.jl:232-260) of row
// i, given its product prod (A dx or A x'); adds dy^2 and Adx*dy.
template <typename T>
__device__ __forceinline__ void dual_update(const DualView<T> &v, int i, const DualRow &r,
                                            double prod, DD &acc_dy, DD &acc_int) {
  constexpr bool kF32 = std::is_same<T, float>::value;
  if (v.pending) {
    v.sy[i] = r.sy + r.y * v.w;
    v.sax[i] = r.sax + r.ax * v.w;
  }
  const double adx = kF32 ? prod - r.ax : prod;
  double t = r.y + v.sig * ((r.b - 2.0 * adx) - r.ax);
  if (i >= v.neq) t = jl_max(t, r.ylo);
  if constexpr (!kF32) {
    const double d = t - r.y;
    v.yn[i] = r.y + d;
    v.axn[i] = r.ax + adx;
    dd_add_prod(acc_dy, d, d);
    dd_add_prod(acc_int, adx, d);
  } else {
    const T yr = (T)t;
    const double d = (double)yr - r.y;
    v.yn[i] = yr;
    v.axn[i] = prod;
    dd_add_prod(acc_dy, d, d);
    dd_add_prod(acc_int, adx, d);
  }
}

// Dual half-step over the slices of A: the product, the pending weighted-sum
// update of y and A x, the projected dual step and the partial sums of dy^2
// and Adx*dy of each slice (part[4p..4p+3] for its partial index p).
template <typename T, int U = 8>
__device__ __forceinline__ void dual_slice(const SliceDesc &sd, int lane,
                                           const InstDev *__restrict__ insts, const Ctl *ctl,
                                           double *part) {
  if (sd.nrows < 0) return;
  const Ctl *C = &ctl[sd.inst];
  const int active = C->active;
  const InstDev &I = insts[sd.inst];
  const DualView<T> v(I, C);
  const int32_t *scol = I.as_col;
  const T *sval = (const T *)I.as_val;
  // as in primal_slice
  asm volatile("" ::"s"(active), "s"(v.yv), "s"(v.bv), "s"(v.dxv), "s"(v.yn), "s"(v.axv),
               "s"(v.axn), "s"(v.sy), "s"(v.sax), "s"(scol), "s"(sval), "s"(C->eta),
               "s"(C->omega), "s"(C->w_pending), "s"(C->pending), "s"(I.neq));
  if (active == 0) return;
  const int i = sd.row0 + lane;
  const bool mine = lane < sd.nrows;
  DualRow r{};
  if (mine) r = dual_load(v, i);
  double prod = group_sum(slice_lane_sum<T, T, U>(scol, sval, v.dxv, sd.off, sd.width, lane), sd.lg);
  prod = __shfl(prod, (lane << sd.lg) & 63, 64);
  DD acc_dy = dd_zero(), acc_int = dd_zero();
  if (mine) dual_update(v, i, r, prod, acc_dy, acc_int);
  acc_dy = wave_dd(acc_dy);
  acc_int = wave_dd(acc_int);
  if (lane == 0) {
    part[4 * sd.part] = acc_dy.s;
    part[4 * sd.part + 1] = acc_dy.e;
    part[4 * sd.part + 2] = acc_int.s;
    part[4 * sd.part + 3] = acc_int.e;
  }
}

template <typename T, int U, int W>
__global__ __launch_bounds__(64 * W) void k_dual(const SliceDesc *__restrict__ slices, int nsl,
                                                 const InstDev *__restrict__ insts,
                                                 const Ctl *__restrict__ ctl,
                                                 double *__restrict__ part) {
  // wave-uniform: the descriptor is one scalar load, and the wave's
  // instance-wide values live in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane((int)blockIdx.x * W + ((int)threadIdx.x >> 6));
  if (wv >= nsl) return;
  dual_slice<T, U>(load_slice(slices, wv), threadIdx.x & 63, insts, ctl, part);
}

// Chunk slices of long rows of A (see k_primal_long).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dual_long(const SliceDesc *__restrict__ slices,
                                                        int nsl, const LongRow *__restrict__ longs,
                                                        int32_t *__restrict__ long_cnt,
                                                        double *__restrict__ long_part,
                                                        const InstDev *__restrict__ insts,
                                                        const Ctl *__restrict__ ctl,
                                                        double *__restrict__ part) {
  const int wv = (int)blockIdx.x * (kThreads / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = slices[wv];
  const Ctl *C = &ctl[sd.inst];
  if (C->active == 0) return;
  const InstDev &I = insts[sd.inst];
  const DualView<T> v(I, C);
  const DD t = wave_dd(slice_lane_sum(I.as_col, (const T *)I.as_val, v.dxv, sd.off, sd.width, lane));
  DD acc_dy = dd_zero(), acc_int = dd_zero();
  if (lane == 0) {
    double tot;
    if (longrow_publish(longs[sd.lid], sd.chunk, sd.lid, long_cnt, long_part, t, &tot))
      dual_update(v, sd.row0, dual_load(v, sd.row0), tot, acc_dy, acc_int);
    part[4 * sd.part] = acc_dy.s;
    part[4 * sd.part + 1] = acc_dy.e;
    part[4 * sd.part + 2] = acc_int.s;
    part[4 * sd.part + 3] = acc_int.e;
  }
}

// Reduction of the step partials in a fixed order and the step rule; one
// work-group of kStepThreads per instance.  Eight waves rather than sixteen:
// with other streams' step kernels on the device, a smaller work-group finds
// a CU with room sooner, and this launch sits on every attempt's critical
// path (512-LP batch in FP64: 50.4 vs 47.6 solves/s; 4 waves 50.2, 1 wave
// 47.8).
constexpr int kStepThreads = 512;

// take_step! with ConstantStepsizeParams (src/primal_dual_hybrid_gradient_gpu.jl:418-452):
// the attempt is always taken with the fixed step size; one KKT pass, no
// attempt count (the reference counts attempts in the adaptive rule only).
__device__ __forceinline__ void step_constant(Ctl *C) {
  C->kkt_passes += 1.0;
  C->cur ^= 1;
  C->pending = 1;
  C->w_pending = C->eta_entry;
  C->wsum += C->eta_entry;
  C->count += 1;
  C->accepted += 1;
  C->aty_valid = 0;
  if (C->accepted >= C->target) C->active = 0;
}

// This thread's share of the partials [b0, b1) of one kernel (S = 2: one
// compensated pair per partial, into a; S = 4: two, into a and c), merged
// in the order b = b0 + thread, + kStepThreads, ...  While eight more
// remain, their loads are issued together, so a long range costs an eighth
// of the memory round trips of one load per merge (the n = 1M LP: 31 primal
// partials per thread); two to seven left are loaded together as well (the
// slots past b1 re-read the last partial and are left out of the merge by a
// select), and a single one alone.
template <int S>
__device__ __forceinline__ void merge_partial(const double *v, DD *a, DD *c) {
  dd_merge(*a, DD{v[0], v[1]});
  if constexpr (S == 4) dd_merge(*c, DD{v[2], v[3]});
}
template <int S>
__device__ __forceinline__ void step_partials(gptr<double> part, int b0, int b1, DD *a, DD *c) {
  constexpr int G = 8;
  int b = b0 + (int)threadIdx.x;
  for (; b + (G - 1) * kStepThreads < b1; b += G * kStepThreads) {
    double v[G][S];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int k = 0; k < S; ++k) v[g][k] = part[S * (b + g * kStepThreads) + k];
#pragma unroll
    for (int g = 0; g < G; ++g) merge_partial<S>(v[g], a, c);
  }
  if (b + kStepThreads < b1) {
    double v[G][S];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = min(b + g * kStepThreads, b1 - 1);
#pragma unroll
      for (int k = 0; k < S; ++k) v[g][k] = part[S * q + k];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      DD ta = *a, tc = c ? *c : dd_zero();
      merge_partial<S>(v[g], &ta, &tc);
      if (b + g * kStepThreads < b1) {
        *a = ta;
        if constexpr (S == 4) *c = tc;
      }
    }
  } else if (b < b1) {
    double v[S];
#pragma unroll
    for (int k = 0; k < S; ++k) v[k] = part[S * b + k];
    merge_partial<S>(v, a, c);
  }
}

// The step of instance q by a work-group of kStepThreads threads; sc is
// LDS scratch of 3 x 2 kStepThreads/64 doubles.  Contains a barrier.
__device__ __forceinline__ void step_instance(int q, const InstDev *__restrict__ insts, Ctl *ctl,
                                              const double *ppart, const double *dpart,
                                              const double *__restrict__ red_tab,
                                              const double *__restrict__ grow_tab,
                                              int64_t tab_len, double red_exp, double grow_exp,
                                              int constant, int64_t debug_attempt,
                                              double (*sc)[2 * kStepThreads / 64]) {
  Ctl *C = &ctl[q];
  // The control block and the instance's partial ranges are independent
  // loads: issue them together, then the step factors of the next attempt
  // beside the partials, so the launch waits for two memory round trips
  // before its reduction rather than four.
  const gptr<Ctl> cg = gp(ctl) + q;
  const gptr<InstDev> ig = gp(insts) + q;
  const int active = cg->active;
  const int64_t next = cg->attempts + 1;
  int ps0 = ig->ps0, pl0 = ig->pl0, ds0 = ig->ds0, dl0 = ig->dl0;
  int ps1 = ig->ps1, pl1 = ig->pl1, ds1 = ig->ds1, dl1 = ig->dl1;
  // An empty asm that takes every value keeps the compiler from sinking the
  // range loads below the test of `active` (it would issue them one round
  // trip later); it holds no instruction.
  asm volatile("" ::"s"(active), "s"(ps0), "s"(ps1), "s"(pl0), "s"(pl1), "s"(ds0), "s"(ds1),
               "s"(dl0), "s"(dl1));
  const bool run = active && !constant;
  if (!run) ps1 = ps0, pl1 = pl0, ds1 = ds0, dl1 = dl0;
  double fr = 0.0, fg = 0.0;
  if (run && next < tab_len) {  // uniform: scalar loads, waited for at their use
    fr = gp(red_tab)[next];
    fg = gp(grow_tab)[next];
  }
  const gptr<double> pp = gp(ppart), dp = gp(dpart);
  DD sx = dd_zero(), sy = dd_zero(), sxy = dd_zero();
  step_partials<2>(pp, ps0, ps1, &sx, nullptr);
  step_partials<2>(pp, pl0, pl1, &sx, nullptr);
  step_partials<4>(dp, ds0, ds1, &sy, &sxy);
  step_partials<4>(dp, dl0, dl1, &sy, &sxy);
  if (!active) return;
  if (constant) {
    if (threadIdx.x == 0) step_constant(C);
    return;
  }
  sx = wave_dd(sx);
  sy = wave_dd(sy);
  sxy = wave_dd(sxy);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sc[0][2 * w] = sx.s; sc[0][2 * w + 1] = sx.e;
    sc[1][2 * w] = sy.s; sc[1][2 * w + 1] = sy.e;
    sc[2][2 * w] = sxy.s; sc[2][2 * w + 1] = sxy.e;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  DD r[3];
  for (int v = 0; v < 3; ++v) {
    r[v] = DD{sc[v][0], sc[v][1]};
    for (int k = 1; k < kStepThreads / 64; ++k) dd_merge(r[v], DD{sc[v][2 * k], sc[v][2 * k + 1]});
  }
  step_rule(C, dd_val(r[0]), dd_val(r[1]), dd_val(r[2]), fr, fg, tab_len, red_exp, grow_exp,
            debug_attempt);
}

__global__ __launch_bounds__(kStepThreads) void k_step(const InstDev *__restrict__ insts,
                                                       Ctl *__restrict__ ctl,
                                                       const double *__restrict__ ppart,
                                                       const double *__restrict__ dpart,
                                                       const double *__restrict__ red_tab,
                                                       const double *__restrict__ grow_tab,
                                                       int64_t tab_len, double red_exp,
                                                       double grow_exp, int constant,
                                                       int64_t debug_attempt) {
  __shared__ double sc[3][2 * kStepThreads / 64];
  step_instance(blockIdx.x, insts, ctl, ppart, dpart, red_tab, grow_tab, tab_len, red_exp,
                grow_exp, constant, debug_attempt, sc);
}

// ---------------------------------------------------------------------------
// Evaluation: averages, unscaled KKT statistics of the average
// (src/iteration_stats_utils_gpu.jl:26-502), the three weighted-KKT
// candidates of the restart scheme (src/saddle_point_gpu.jl:250-278, 479-622)
// and the restart distances (:673-682), all in one pass.  Blocks [0, np) walk
// the A^T row blocks (variables), the rest walk the A row blocks (rows).
// Each accumulator is a compensated sum or an exact maximum; partials hold
// (s, e) pairs per block.

// Row blocks per work-group of k_eval: its 25 / 22 compensated accumulators
// are reduced once per run of an instance's blocks in the group rather than
// once per block (a row block holds about 128 rows of a random LP).
constexpr int kEvalGroup = 8;

constexpr uint32_t kEvalMaxN = (1u << 1) | (1u << 3) | (1u << 6) | (1u << 8) | (1u << 10) | (1u << 11);
constexpr uint32_t kEvalMaxM = (1u << 1) | (1u << 3) | (1u << 5) | (1u << 7);
__device__ __forceinline__ bool eval_is_max_n(int k) { return (kEvalMaxN >> k) & 1u; }
__device__ __forceinline__ bool eval_is_max_m(int k) { return (kEvalMaxM >> k) & 1u; }

__device__ __forceinline__ double contrib(double l, double u, double r) {
  return r > 0.0 ? l * r : (r < 0.0 ? u * r : 0.0);
}

template <int K>
__device__ __forceinline__ void block_reduce_acc(DD (&acc)[K], uint32_t maxmask, double *lds,
                                                 double *out) {
  const int w = threadIdx.x >> 6;
  constexpr int NW = kThreads / 64;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool ismax = (maxmask >> k) & 1u;
    DD v;
    if (ismax) {
      v = DD{wave_max(acc[k].s), 0.0};
    } else {
      v = wave_dd(acc[k]);
    }
    if ((threadIdx.x & 63) == 0) {
      lds[2 * (k * NW + w)] = v.s;
      lds[2 * (k * NW + w) + 1] = v.e;
    }
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const bool ismax = (maxmask >> k) & 1u;
    DD r{lds[2 * (k * NW)], lds[2 * (k * NW) + 1]};
    for (int q = 1; q < NW; ++q) {
      const DD v{lds[2 * (k * NW + q)], lds[2 * (k * NW + q) + 1]};
      if (ismax)
        r.s = jl_max(r.s, v.s);
      else
        dd_merge(r, v);
    }
    out[2 * k] = r.s;
    out[2 * k + 1] = r.e;
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

// Row blocks whose rows k_eval summed into the partial of an earlier block
// of the same instance get zero partials (k_eval_finish adds them exactly).
__device__ __forceinline__ void zero_partials(double *ep, int b0, int b1, int K) {
  double *p = ep + (size_t)b0 * 2 * K;
  const int count = (b1 - b0) * 2 * K;
  for (int i = (int)threadIdx.x; i < count; i += kThreads) p[i] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_eval(const BlockDesc *__restrict__ pblocks, int np,
                                                   const BlockDesc *__restrict__ dblocks, int nd,
                                                   const InstDev *__restrict__ insts,
                                                   const Ctl *__restrict__ ctl,
                                                   double *__restrict__ epn,
                                                   double *__restrict__ epm) {
  __shared__ double lds[2 * kEvalN * (kThreads / 64)];
  const int nsp = (np + kEvalGroup - 1) / kEvalGroup;
  const bool nside = (int)blockIdx.x < nsp;
  const int sb = nside ? (int)blockIdx.x : (int)blockIdx.x - nsp;
  const BlockDesc *__restrict__ blocks = nside ? pblocks : dblocks;
  const int bend = min(nside ? np : nd, (sb + 1) * kEvalGroup);
  for (int b0 = sb * kEvalGroup; b0 < bend;) {
    // the run [bfirst, blast) of this group's row blocks that belong to one instance
    const int bfirst = b0;
    const int qi = blocks[bfirst].inst;
    int blast = bfirst + 1;
    while (blast < bend && blocks[blast].inst == qi) ++blast;
    b0 = blast;
    const Ctl *C = &ctl[qi];
    if (!C->eval_flag) continue;
    const InstDev &I = insts[qi];
    const int cur = C->cur;
    const bool usecur = C->use_current != 0;
    const double wsum = C->wsum;
    if (nside) {
      DD acc[kEvalN];
#pragma unroll
      for (int k = 0; k < kEvalN; ++k) acc[k] = dd_zero();
      const void *xcur = cur ? I.x1 : I.x0;
      for (int bb = bfirst; bb < blast; ++bb) {
        int r0, r1;
        block_rows(blocks[bb], &r0, &r1);
        for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
          const double xc = ld<T>(xcur, j);
          const double cs = ld<T>(I.c, j), ls = ld<T>(I.l, j), us = ld<T>(I.u, j);
          const double ax = usecur ? xc : ld<double>(I.sx, j) / wsum;
          const double ag = usecur ? ld<T>(I.gbuf, j) : (-ld<double>(I.saty, j)) / wsum + cs;
          const double gnew = cs - ld<T>(I.aty, j);
          st<T>(I.avg_x, j, ax);
          st<T>(I.avg_g, j, ag);
          st<T>(I.gbuf, j, gnew);
          // original space
          const double Dj = gp(I.D)[j];
          const double ox = ax / Dj, og = ag * Dj;
          const double ol = gp(I.ol)[j], ou = gp(I.ou)[j], oc = gp(I.oc)[j];
          const double fl = isfinite(ol) ? 1.0 : 0.0, fu = isfinite(ou) ? 1.0 : 0.0;
          const double lvv = jl_max(ol - ox, 0.0), uvv = jl_max(ox - ou, 0.0);
          dd_add_prod(acc[0], lvv, lvv);
          dd_add_prod(acc[0], uvv, uvv);
          acc[1].s = jl_max(acc[1].s, jl_max(lvv, uvv));
          dd_add_prod(acc[2], oc, ox);
          acc[3].s = jl_max(acc[3].s, fabs(ox));
          dd_add_prod(acc[4], ox, ox);
          const double r = jl_max(og, 0.0) * fl + jl_min(og, 0.0) * fu;
          const double rv = og - r;
          dd_add_prod(acc[5], rv, rv);
          acc[6].s = jl_max(acc[6].s, fabs(rv));
          dd_add(acc[7], contrib(ol, ou, r));
          const double g2 = og - oc;
          const double r2 = jl_max(g2, 0.0) * fl + jl_min(g2, 0.0) * fu;
          const double rv2 = g2 - r2;
          acc[8].s = jl_max(acc[8].s, fabs(rv2));
          dd_add(acc[9], contrib(ol, ou, r2));
          acc[10].s = jl_max(acc[10].s, fabs(r2));
          acc[11].s = jl_max(acc[11].s, jl_max(fl != 0.0 ? jl_max(-ox, 0.0) : 0.0,
                                                fu != 0.0 ? jl_max(ox, 0.0) : 0.0));
          // restart candidates in the scaled space: current, average, last restart
          const double xs3[3] = {xc, ax, ld<T>(I.lx, j)};
          const double gs3[3] = {gnew, ag, ld<T>(I.lg, j)};
          const double sfl = isfinite(ls) ? 1.0 : 0.0, sfu = isfinite(us) ? 1.0 : 0.0;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double xs = xs3[q], gs = gs3[q];
            const double a = jl_max(ls - xs, 0.0), b = jl_max(xs - us, 0.0);
            dd_add_prod(acc[12 + 4 * q], a, a);
            dd_add_prod(acc[12 + 4 * q], b, b);
            dd_add_prod(acc[13 + 4 * q], cs, xs);
            const double rr = jl_max(gs, 0.0) * sfl + jl_min(gs, 0.0) * sfu;
            const double rvv = gs - rr;
            dd_add_prod(acc[14 + 4 * q], rvv, rvv);
            dd_add(acc[15 + 4 * q], contrib(ls, us, rr));
          }
          const double dxl = ax - xs3[2];
          dd_add_prod(acc[24], dxl, dxl);
        }
      }
      block_reduce_acc<kEvalN>(acc, kEvalMaxN, lds, epn + (size_t)bfirst * 2 * kEvalN);
      zero_partials(epn, bfirst + 1, blast, kEvalN);
    } else {
      DD acc[kEvalM];
#pragma unroll
      for (int k = 0; k < kEvalM; ++k) acc[k] = dd_zero();
      const void *ycur = cur ? I.y1 : I.y0;
      const void *axcur = cur ? I.ax1 : I.ax0;
      const int neq = I.neq;
      const bool has_ylo = I.ylo != nullptr;
      for (int bb = bfirst; bb < blast; ++bb) {
        int r0, r1;
        block_rows(blocks[bb], &r0, &r1);
        for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
          const double yc = ld<T>(ycur, i), axc = ld<double>(axcur, i);
          const double ay = usecur ? yc : ld<double>(I.sy, i) / wsum;
          const double aax = usecur ? axc : ld<double>(I.sax, i) / wsum;
          st<T>(I.avg_y, i, ay);
          st<T>(I.avg_ax, i, aax);
          const double Ei = gp(I.E)[i];
          const double oy = ay / Ei, oax = aax * Ei;
          const double ob = gp(I.ob)[i];
          const double t = ob - oax;
          const double cvv = i < neq ? t : jl_max(t, 0.0);
          dd_add_prod(acc[0], cvv, cvv);
          acc[1].s = jl_max(acc[1].s, fabs(cvv));
          if (i >= neq) {
            // dual residual max(ylo - y, 0); with the primal-objective term
            // -ylo * max(Ax - b, 0) of a shifted dual bound (refinement only)
            const double yl = has_ylo ? gp(I.oylo)[i] : 0.0;
            const double dr = has_ylo ? jl_max(yl - oy, 0.0) : jl_max(-oy, 0.0);
            dd_add_prod(acc[2], dr, dr);
            acc[3].s = jl_max(acc[3].s, dr);
            if (has_ylo) dd_add_prod(acc[18], yl, jl_max(oax - ob, 0.0));
          }
          dd_add_prod(acc[4], ob, oy);
          acc[5].s = jl_max(acc[5].s, fabs(oy));
          dd_add_prod(acc[6], oy, oy);
          acc[7].s = jl_max(acc[7].s, i < neq ? fabs(oax) : jl_max(-oax, 0.0));
          const double bs = ld<T>(I.b, i);
          const double ys3[3] = {yc, ay, ld<T>(I.ly, i)};
          const double as3[3] = {axc, aax, ld<T>(I.lax, i)};
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double tt = bs - as3[q];
            const double cc = i < neq ? tt : jl_max(tt, 0.0);
            dd_add_prod(acc[8 + 3 * q], cc, cc);
            if (i >= neq) {
              const double yl = has_ylo ? ld<T>(I.ylo, i) : 0.0;
              const double dr = has_ylo ? jl_max(yl - ys3[q], 0.0) : jl_max(-ys3[q], 0.0);
              dd_add_prod(acc[9 + 3 * q], dr, dr);
              if (has_ylo) dd_add_prod(acc[19 + q], yl, jl_max(as3[q] - bs, 0.0));
            }
            dd_add_prod(acc[10 + 3 * q], bs, ys3[q]);
          }
          const double dyl = ay - ys3[2];
          dd_add_prod(acc[17], dyl, dyl);
        }
      }
      block_reduce_acc<kEvalM>(acc, kEvalMaxM, lds, epm + (size_t)bfirst * 2 * kEvalM);
      zero_partials(epm, bfirst + 1, blast, kEvalM);
    }
    __syncthreads();  // lds is reused by the next run
  }
}

// Per-instance combination of block partials: one work-group per
// (instance, accumulator), blocks combined in a fixed strided order.
// Accumulator k < KN comes from epn, the rest from epm; bit k of maxN /
// bit (k - KN) of maxM marks a maximum (else a compensated sum).
__global__ __launch_bounds__(kThreads) void k_eval_finish(const InstDev *__restrict__ insts,
                                                          const Ctl *__restrict__ ctl,
                                                          const double *__restrict__ epn,
                                                          const double *__restrict__ epm,
                                                          double *__restrict__ out, int KN,
                                                          int KM, uint32_t maxN, uint32_t maxM,
                                                          int stride) {
  __shared__ double scratch[2 * kThreads / 64];
  const int q = blockIdx.x;
  const int k = blockIdx.y;
  if (!ctl[q].eval_flag) return;
  const InstDev &I = insts[q];
  const bool nside = k < KN;
  const int kk = nside ? k : k - KN;
  const int K = nside ? KN : KM;
  const bool ismax = ((nside ? maxN : maxM) >> kk) & 1u;
  const double *ep = nside ? epn : epm;
  const int b0 = nside ? I.pb0 : I.db0, b1 = nside ? I.pb1 : I.db1;
  if (ismax) {
    double s = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += kThreads)
      s = jl_max(s, ep[((size_t)b * K + kk) * 2]);
    s = wave_max(s);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scratch[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = scratch[0];
      for (int i = 1; i < kThreads / 64; ++i) t = jl_max(t, scratch[i]);
      out[(size_t)q * stride + k] = t;
    }
  } else {
    DD s = dd_zero();
    for (int b = b0 + (int)threadIdx.x; b < b1; b += kThreads)
      dd_merge(s, DD{ep[((size_t)b * K + kk) * 2], ep[((size_t)b * K + kk) * 2 + 1]});
    s = block_dd(s, scratch);
    if (threadIdx.x == 0) out[(size_t)q * stride + k] = dd_val(s);
  }
}

// Second pass for the primal ray objective dot(c, x / norm(x, Inf))
// (src/iteration_stats_utils_gpu.jl:326-343), which needs the norm first.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_eval_ray(const BlockDesc *__restrict__ pblocks,
                                                       const InstDev *__restrict__ insts,
                                                       const Ctl *__restrict__ ctl,
                                                       const double *__restrict__ out,
                                                       double *__restrict__ epr) {
  __shared__ double scratch[2 * kThreads / 64];
  const BlockDesc bd = pblocks[blockIdx.x];
  if (!ctl[bd.inst].eval_flag) return;
  const InstDev &I = insts[bd.inst];
  const double sinf = out[(size_t)bd.inst * kEvalStride + 3];
  int r0, r1;
  block_rows(bd, &r0, &r1);
  DD acc = dd_zero();
  if (sinf != 0.0) {
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      const double ox = ld<T>(I.avg_x, j) / I.D[j];
      dd_add_prod(acc, I.oc[j], ox / sinf);
    }
  }
  acc = block_dd(acc, scratch);
  if (threadIdx.x == 0) {
    epr[2 * blockIdx.x] = acc.s;
    epr[2 * blockIdx.x + 1] = acc.e;
  }
}

__global__ __launch_bounds__(kThreads) void k_eval_ray_finish(const InstDev *__restrict__ insts,
                                                              const Ctl *__restrict__ ctl,
                                                              const double *__restrict__ epr,
                                                              double *__restrict__ out) {
  __shared__ double scratch[2 * kThreads / 64];
  const int q = blockIdx.x;
  if (!ctl[q].eval_flag) return;
  const InstDev &I = insts[q];
  DD s = dd_zero();
  for (int b = I.pb0 + (int)threadIdx.x; b < I.pb1; b += kThreads)
    dd_merge(s, DD{epr[2 * b], epr[2 * b + 1]});
  s = block_dd(s, scratch);
  if (threadIdx.x == 0) out[(size_t)q * kEvalStride + kEvalOut] = dd_val(s);
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
      ((double *)I.sx)[j] = 0.0;
      ((double *)I.saty)[j] = 0.0;
      ((T *)I.lx)[j] = xc[j];
      ((T *)I.lg)[j] = ((const T *)I.gbuf)[j];
    }
  } else {
    T *yc = (T *)(cur ? I.y1 : I.y0);
    double *axc = (double *)(cur ? I.ax1 : I.ax0);
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      if (action == 2) {
        yc[i] = ((const T *)I.avg_y)[i];
        axc[i] = (double)((const T *)I.avg_ax)[i];
      }
      ((double *)I.sy)[i] = 0.0;
      ((double *)I.sax)[i] = 0.0;
      ((T *)I.ly)[i] = yc[i];
      ((T *)I.lax)[i] = (T)axc[i];
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 + FP64 iterative refinement (no reference counterpart: the reference
// iterates in FP64 only).  Round k keeps an FP64 point (xk, yk); the FP32
// engine solves the correction LP
//   min  dD (c - A^T yk)' xh + ...  s.t.  A xh (=, >=) dP (b - A xk),
//        dP (l - xk) <= xh <= dP (u - xk),   yh_i >= -dD yk_i on ">=" rows,
// i.e. the original saddle problem shifted to (xk, yk) and scaled by (dP, dD);
// then xk += xh / dP, yk += yh / dD.  Residuals, statistics and the update
// are FP64 with compensated sums.

constexpr uint32_t kRefMaxN = (1u << 1) | (1u << 3) | (1u << 6);
constexpr uint32_t kRefMaxM = (1u << 1) | (1u << 3) | (1u << 5);

// xk += (avg_x / D) / dP clamped to [l, u]; yk += (avg_y / E) / dD, ">=" rows >= 0.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_ref_update(const BlockDesc *__restrict__ pblocks,
                                                         int np,
                                                         const BlockDesc *__restrict__ dblocks,
                                                         const InstDev *__restrict__ insts,
                                                         const Ctl *__restrict__ ctl) {
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  const Ctl *C = &ctl[bd.inst];
  if (C->ref_flag != 1) return;
  const InstDev &I = insts[bd.inst];
  int r0, r1;
  block_rows(bd, &r0, &r1);
  if (nside) {
    const double dp = C->ref_dp;
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      const double xh = ld<T>(I.avg_x, j) / I.D[j];
      const double x = I.xk[j] + xh / dp;
      I.xk[j] = jl_min(I.ru[j], jl_max(I.rl[j], x));
    }
  } else {
    const double ddl = C->ref_dd;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      const double yh = ld<T>(I.avg_y, i) / I.E[i];
      double y = I.yk[i] + yh / ddl;
      if (i >= I.neq) y = jl_max(y, 0.0);
      I.yk[i] = y;
    }
  }
}

// raty = A^T yk and rax = A xk with the unscaled FP64 matrix.
__global__ __launch_bounds__(kThreads) void k_ref_spmv(const BlockDesc *__restrict__ pblocks,
                                                       int np,
                                                       const BlockDesc *__restrict__ dblocks,
                                                       const LongRow *__restrict__ longs,
                                                       int32_t *__restrict__ long_cnt,
                                                       double *__restrict__ long_part,
                                                       const InstDev *__restrict__ insts,
                                                       const Ctl *__restrict__ ctl) {
  __shared__ double lds_p[kBlockNnz];
  __shared__ double lds_q[kBlockNnz];
  __shared__ double rowsum[kBlockRows];
  __shared__ double scratch[2 * kThreads / 64];
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  if (ctl[bd.inst].ref_flag != 1) return;
  const InstDev &I = insts[bd.inst];
  const int32_t *ptr = nside ? I.t_ptr : I.a_ptr;
  const int32_t *col = nside ? I.t_col : I.a_col;
  const double *val = nside ? I.t_orig : I.a_orig;
  const double *xv = nside ? I.yk : I.xk;
  double *outv = nside ? I.raty : I.rax;
  if (bd.nrows >= 0) {
    rowblock_sums(ptr, col, val, xv, bd.row0, bd.nrows, lds_p, lds_q, rowsum);
    for (int r = threadIdx.x; r < bd.nrows; r += kThreads) outv[bd.row0 + r] = rowsum[r];
  } else {
    const int chunk = -bd.nrows - 1;
    const DD sm = longrow_chunk(ptr, col, val, xv, bd.row0, chunk, scratch);
    if (threadIdx.x == 0) {
      double tot;
      if (longrow_publish(longs[bd.aux], chunk, bd.aux, long_cnt, long_part, sm, &tot))
        outv[bd.row0] = tot;
    }
  }
}

// KKT statistics of (xk, yk) for the true problem, plus the norms the next
// correction LP needs.  N: 0 sum lv^2+uv^2, 1 max viol, 2 c'x, 3 max|x|, 4 sum x^2,
// 5 sum rv^2, 6 max|rv|, 7 sum contributions, 8 sum (rv/D)^2 (dual violation,
// scaled space); M: 0 sum cv^2, 1 max|cv|, 2 sum dres^2, 3 max dres, 4 b'y,
// 5 max|y|, 6 sum y^2, 7 sum (cv/E)^2 (primal violation, scaled space),
// 8 sum (dres E)^2.
__global__ __launch_bounds__(kThreads) void k_ref_stats(const BlockDesc *__restrict__ pblocks,
                                                        int np,
                                                        const BlockDesc *__restrict__ dblocks,
                                                        const InstDev *__restrict__ insts,
                                                        const Ctl *__restrict__ ctl,
                                                        double *__restrict__ epn,
                                                        double *__restrict__ epm) {
  __shared__ double lds[2 * kRefN * (kThreads / 64)];
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  if (ctl[bd.inst].ref_flag != 1) return;
  const InstDev &I = insts[bd.inst];
  int r0, r1;
  block_rows(bd, &r0, &r1);
  if (nside) {
    DD acc[kRefN];
#pragma unroll
    for (int k = 0; k < kRefN; ++k) acc[k] = dd_zero();
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      const double x = I.xk[j], l = I.rl[j], u = I.ru[j], c = I.rc[j];
      const double fl = isfinite(l) ? 1.0 : 0.0, fu = isfinite(u) ? 1.0 : 0.0;
      const double lvv = jl_max(l - x, 0.0), uvv = jl_max(x - u, 0.0);
      dd_add_prod(acc[0], lvv, lvv);
      dd_add_prod(acc[0], uvv, uvv);
      acc[1].s = jl_max(acc[1].s, jl_max(lvv, uvv));
      dd_add_prod(acc[2], c, x);
      acc[3].s = jl_max(acc[3].s, fabs(x));
      dd_add_prod(acc[4], x, x);
      const double g = c - I.raty[j];
      const double r = jl_max(g, 0.0) * fl + jl_min(g, 0.0) * fu;
      const double rv = g - r;
      dd_add_prod(acc[5], rv, rv);
      acc[6].s = jl_max(acc[6].s, fabs(rv));
      dd_add(acc[7], contrib(l, u, r));
      const double rvs = rv / I.D[j];  // the violation in the scaled space
      dd_add_prod(acc[8], rvs, rvs);
    }
    block_reduce_acc<kRefN>(acc, kRefMaxN, lds, epn + (size_t)blk * 2 * kRefN);
  } else {
    DD acc[kRefM];
#pragma unroll
    for (int k = 0; k < kRefM; ++k) acc[k] = dd_zero();
    const int neq = I.neq;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      const double y = I.yk[i], b = I.rb[i];
      const double t = b - I.rax[i];
      const double cvv = i < neq ? t : jl_max(t, 0.0);
      dd_add_prod(acc[0], cvv, cvv);
      acc[1].s = jl_max(acc[1].s, fabs(cvv));
      if (i >= neq) {
        const double dr = jl_max(-y, 0.0);
        dd_add_prod(acc[2], dr, dr);
        acc[3].s = jl_max(acc[3].s, dr);
      }
      dd_add_prod(acc[4], b, y);
      acc[5].s = jl_max(acc[5].s, fabs(y));
      dd_add_prod(acc[6], y, y);
      const double cs = cvv / I.E[i];  // violations in the scaled space
      dd_add_prod(acc[7], cs, cs);
      if (i >= neq) {
        const double ds = jl_max(-y, 0.0) * I.E[i];
        dd_add_prod(acc[8], ds, ds);
      }
    }
    block_reduce_acc<kRefM>(acc, kRefMaxM, lds, epm + (size_t)blk * 2 * kRefM);
  }
}

// Data of the next correction LP (original and scaled) and a fresh PDHG state.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_ref_setup(const BlockDesc *__restrict__ pblocks,
                                                        int np,
                                                        const BlockDesc *__restrict__ dblocks,
                                                        const InstDev *__restrict__ insts,
                                                        const Ctl *__restrict__ ctl) {
  const bool nside = (int)blockIdx.x < np;
  const int blk = nside ? (int)blockIdx.x : (int)blockIdx.x - np;
  const BlockDesc bd = nside ? pblocks[blk] : dblocks[blk];
  const Ctl *C = &ctl[bd.inst];
  if (C->ref_flag != 2) return;
  const InstDev &I = insts[bd.inst];
  const double dp = C->ref_dp, ddl = C->ref_dd;
  int r0, r1;
  block_rows(bd, &r0, &r1);
  const bool keep = C->ref_keep != 0;  // (xk, yk) is the best refined point so far
  if (nside) {
    for (int j = r0 + (int)threadIdx.x; j < r1; j += kThreads) {
      const double x = I.xk[j];
      if (keep) I.xb[j] = x;
      const double ch = ddl * (I.rc[j] - I.raty[j]);
      const double lh = dp * (I.rl[j] - x), uh = dp * (I.ru[j] - x);
      ((double *)I.oc)[j] = ch;
      ((double *)I.ol)[j] = lh;
      ((double *)I.ou)[j] = uh;
      const double Dj = I.D[j];
      const double cs = ch / Dj;
      st<T>((void *)I.c, j, cs);
      st<T>((void *)I.l, j, lh * Dj);
      st<T>((void *)I.u, j, uh * Dj);
      st<T>(I.gbuf, j, cs);
      st<T>(I.lg, j, cs);
      void *zero[] = {I.x0, I.x1, I.aty, I.dx, I.avg_x, I.avg_g, I.lx};
#pragma unroll
      for (int k = 0; k < 7; ++k) st<T>(zero[k], j, 0.0);
      ((double *)I.sx)[j] = 0.0;
      ((double *)I.saty)[j] = 0.0;
    }
  } else {
    const int neq = I.neq;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += kThreads) {
      if (keep) I.yb[i] = I.yk[i];
      const double bh = dp * (I.rb[i] - I.rax[i]);
      ((double *)I.ob)[i] = bh;
      const double Ei = I.E[i];
      st<T>((void *)I.b, i, bh / Ei);
      const double yl = i >= neq ? -ddl * I.yk[i] : 0.0;
      ((double *)I.oylo)[i] = yl;
      st<T>(I.ylo, i, yl * Ei);
      void *zero[] = {I.y0, I.y1, I.avg_y, I.avg_ax, I.ly, I.lax};
#pragma unroll
      for (int k = 0; k < 6; ++k) st<T>(zero[k], i, 0.0);
      void *zero64[] = {I.ax0, I.ax1, I.sy, I.sax};
#pragma unroll
      for (int k = 0; k < 4; ++k) ((double *)zero64[k])[i] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers.

static inline unsigned grid_of(int n) { return (unsigned)(n > 0 ? n : 0); }
static inline int eval_groups(int nblocks) { return (nblocks + kEvalGroup - 1) / kEvalGroup; }

#define PDLP_CHECK_LAUNCH() return (int)hipGetLastError()

int launch_primal_flush(hipStream_t s, int prec, int nblocks, const BlockDesc *blocks,
                        const LongRow *longs, int32_t *long_cnt, double *long_part,
                        const InstDev *insts, const Ctl *ctl) {
  if (nblocks <= 0) return 0;
  dim3 g(grid_of(nblocks)), b(kThreads);
  if (prec == 0)
    hipLaunchKernelGGL(k_primal_flush<double>, g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl);
  else
    hipLaunchKernelGGL(k_primal_flush<float>, g, b, 0, s, blocks, longs, long_cnt, long_part, insts, ctl);
  PDLP_CHECK_LAUNCH();
}

int launch_dual_flush(hipStream_t s, int prec, int nblocks, const BlockDesc *blocks,
                      const InstDev *insts, const Ctl *ctl) {
  if (nblocks <= 0) return 0;
  dim3 g(grid_of(nblocks)), b(kThreads);
  if (prec == 0)
    hipLaunchKernelGGL(k_dual_flush<double>, g, b, 0, s, blocks, insts, ctl);
  else
    hipLaunchKernelGGL(k_dual_flush<float>, g, b, 0, s, blocks, insts, ctl);
  PDLP_CHECK_LAUNCH();
}

static inline unsigned grid_of_slices(int nsl) {
  return (unsigned)(nsl > 0 ? (nsl + kThreads / 64 - 1) / (kThreads / 64) : 0);
}

// A launch that, when e0 / e1 are given, records them at the start / end of
// the kernel's own execution (hipExtLaunchKernelGGL), so profiled durations
// exclude the dispatch gaps between kernels.
template <typename F, typename... Args>
static void launch_ev(F kernel, dim3 g, dim3 b, unsigned shm, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1, Args... args) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kernel, g, b, shm, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, g, b, shm, s, args...);
}

unsigned slice_lds_bytes(int occupancy, int waves, int lds_per_cu, int lds_per_block) {
  if (occupancy <= 0 || lds_per_cu <= 0) return 0;
  const int w = waves == 1 || waves == 2 ? waves : 4;
  // one wave of a work-group per SIMD (four SIMDs per CU): at most
  // occupancy * 4 / w work-groups of w waves keep occupancy waves per SIMD
  const int groups = std::max(1, occupancy * 4 / w);
  unsigned bytes = (unsigned)(lds_per_cu / groups) & ~1023u;
  if (lds_per_block > 0 && bytes > (unsigned)lds_per_block) bytes = (unsigned)lds_per_block & ~1023u;
  return bytes;
}

// The slice-kernel variant: U in {4, 8, 16} loads in flight per lane, W in
// {1, 2, 4} waves per work-group (anything else: 4).  unroll 0 takes each
// kernel's default: 8, except 16 for the FP64 dual kernel, whose rows (A's,
// about twice as long as A^T's on the LPs measured) then issue a whole row's
// gathers at once; at its default cap of 3 waves per SIMD its 103 VGPRs cost
// no occupancy.  The FP32 dual kernel and both primal kernels are slower with
// 16.  The variants compute the same sums in the same order.
using SliceKernel = void (*)(const SliceDesc *, int, const InstDev *, const Ctl *, double *);
template <typename T, int U>
static SliceKernel primal_variant(int w) {
  return w == 1 ? &k_primal<T, U, 1> : w == 2 ? &k_primal<T, U, 2> : &k_primal<T, U, 4>;
}
template <typename T, int U>
static SliceKernel dual_variant(int w) {
  return w == 1 ? &k_dual<T, U, 1> : w == 2 ? &k_dual<T, U, 2> : &k_dual<T, U, 4>;
}
static int variant_waves(int waves) { return waves == 1 || waves == 2 ? waves : 4; }
static SliceKernel primal_kernel(int prec, int unroll, int waves) {
  const int w = variant_waves(waves);
  if (prec == 0)
    return unroll == 4    ? primal_variant<double, 4>(w)
           : unroll == 16 ? primal_variant<double, 16>(w)
                          : primal_variant<double, 8>(w);
  return unroll == 4    ? primal_variant<float, 4>(w)
         : unroll == 16 ? primal_variant<float, 16>(w)
                        : primal_variant<float, 8>(w);
}
static SliceKernel dual_kernel(int prec, int unroll, int waves) {
  const int w = variant_waves(waves);
  if (prec == 0)
    return unroll == 4   ? dual_variant<double, 4>(w)
           : unroll == 8 ? dual_variant<double, 8>(w)
                         : dual_variant<double, 16>(w);
  return unroll == 4    ? dual_variant<float, 4>(w)
         : unroll == 16 ? dual_variant<float, 16>(w)
                        : dual_variant<float, 8>(w);
}
static inline unsigned grid_of_waves(int nsl, int w) {
  return (unsigned)(nsl > 0 ? (nsl + w - 1) / w : 0);
}

// Normal slices, then the chunk slices of long rows; each slice writes its
// partial sums at its own partial index.  e0 / e1 (optional) bracket both
// launches.
int launch_primal(hipStream_t s, int prec, int nslices, const SliceDesc *slices, int nlong,
                  const SliceDesc *lslices, const LongRow *longs, int32_t *long_cnt,
                  double *long_part, const InstDev *insts, const Ctl *ctl, double *part,
                  int unroll, int waves, unsigned lds_bytes, hipEvent_t e0, hipEvent_t e1) {
  const dim3 b(kThreads);
  if (nslices <= 0 && nlong <= 0) {
    hipError_t r = e0 ? hipEventRecord(e0, s) : hipSuccess;
    if (r == hipSuccess && e1) r = hipEventRecord(e1, s);
    return (int)r;
  }
  if (nslices > 0) {
    hipEvent_t stop = nlong > 0 ? nullptr : e1;
    const int w = variant_waves(waves);
    launch_ev(primal_kernel(prec, unroll, waves), dim3(grid_of_waves(nslices, w)), dim3(64 * w),
              lds_bytes, s, e0, stop, slices, nslices, insts, ctl, part);
  }
  if (nlong > 0) {
    hipEvent_t start = nslices > 0 ? nullptr : e0;
    if (prec == 0)
      launch_ev(k_primal_long<double>, dim3(grid_of_slices(nlong)), b, 0, s, start, e1, lslices, nlong, longs, long_cnt, long_part, insts, ctl, part);
    else
      launch_ev(k_primal_long<float>, dim3(grid_of_slices(nlong)), b, 0, s, start, e1, lslices, nlong, longs, long_cnt, long_part, insts, ctl, part);
  }
  PDLP_CHECK_LAUNCH();
}

int launch_dual(hipStream_t s, int prec, int nslices, const SliceDesc *slices, int nlong,
                const SliceDesc *lslices, const LongRow *longs, int32_t *long_cnt,
                double *long_part, const InstDev *insts, const Ctl *ctl, double *part,
                int unroll, int waves, unsigned lds_bytes, hipEvent_t e0, hipEvent_t e1) {
  const dim3 b(kThreads);
  if (nslices <= 0 && nlong <= 0) {
    hipError_t r = e0 ? hipEventRecord(e0, s) : hipSuccess;
    if (r == hipSuccess && e1) r = hipEventRecord(e1, s);
    return (int)r;
  }
  if (nslices > 0) {
    hipEvent_t stop = nlong > 0 ? nullptr : e1;
    const int w = variant_waves(waves);
    launch_ev(dual_kernel(prec, unroll, waves), dim3(grid_of_waves(nslices, w)), dim3(64 * w),
              lds_bytes, s, e0, stop, slices, nslices, insts, ctl, part);
  }
  if (nlong > 0) {
    hipEvent_t start = nslices > 0 ? nullptr : e0;
    if (prec == 0)
      launch_ev(k_dual_long<double>, dim3(grid_of_slices(nlong)), b, 0, s, start, e1, lslices, nlong, longs, long_cnt, long_part, insts, ctl, part);
    else
      launch_ev(k_dual_long<float>, dim3(grid_of_slices(nlong)), b, 0, s, start, e1, lslices, nlong, longs, long_cnt, long_part, insts, ctl, part);
  }
  PDLP_CHECK_LAUNCH();
}

// Sliced copy of a CSR matrix (one wave per slice; see SliceDesc).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_slice_fill(const SliceDesc *__restrict__ slices,
                                                         int nsl, const int32_t *__restrict__ ptr,
                                                         const int32_t *__restrict__ col,
                                                         const T *__restrict__ val,
                                                         int32_t *__restrict__ scol,
                                                         T *__restrict__ sval) {
  const int wv = (int)blockIdx.x * (kThreads / 64) + ((int)threadIdx.x >> 6);
  if (wv >= nsl) return;
  const int lane = threadIdx.x & 63;
  const SliceDesc sd = slices[wv];
  if (sd.nrows < 0) return;
  int row, k, kstep, len;
  if (sd.nrows == 0) {
    row = sd.row0;
    k = sd.chunk * kSellChunk + lane;
    kstep = 64;
    len = ptr[row + 1] - ptr[row];
  } else {
    const int g = lane >> sd.lg;
    row = sd.row0 + g;
    k = lane & ((1 << sd.lg) - 1);
    kstep = 1 << sd.lg;
    len = g < sd.nrows ? ptr[row + 1] - ptr[row] : 0;
  }
  const int base = len > 0 ? ptr[row] : 0;
  for (int q = 0; q < sd.width; ++q, k += kstep) {
    const int slot = sd.off + 64 * q + lane;
    const bool ok = k < len;
    scol[slot] = ok ? col[base + k] : -1;
    sval[slot] = ok ? val[base + k] : (T)0;
  }
}

int launch_slice_fill(hipStream_t s, int prec, int nslices, const SliceDesc *slices,
                      const int32_t *ptr, const int32_t *col, const void *val, int32_t *scol,
                      void *sval) {
  if (nslices <= 0) return 0;
  dim3 g(grid_of_slices(nslices)), b(kThreads);
  if (prec == 0)
    hipLaunchKernelGGL(k_slice_fill<double>, g, b, 0, s, slices, nslices, ptr, col,
                       (const double *)val, scol, (double *)sval);
  else
    hipLaunchKernelGGL(k_slice_fill<float>, g, b, 0, s, slices, nslices, ptr, col,
                       (const float *)val, scol, (float *)sval);
  PDLP_CHECK_LAUNCH();
}

int launch_step(hipStream_t s, int ninst, const InstDev *insts, Ctl *ctl, const double *ppart,
                const double *dpart, const double *red_tab, const double *grow_tab,
                int64_t tab_len, double red_exp, double grow_exp, int constant,
                int64_t debug_attempt) {
  if (ninst <= 0) return 0;
  hipLaunchKernelGGL(k_step, dim3(grid_of(ninst)), dim3(kStepThreads), 0, s, insts, ctl, ppart,
                     dpart, red_tab, grow_tab, tab_len, red_exp, grow_exp, constant,
                     debug_attempt);
  PDLP_CHECK_LAUNCH();
}

int launch_eval(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl, double *epn,
                double *epm, double *epr, int ninst, double *out) {
  if (np + nd > 0) {
    if (prec == 0)
      hipLaunchKernelGGL(k_eval<double>, dim3(grid_of(eval_groups(np) + eval_groups(nd))),
                         dim3(kThreads), 0, s, pblocks, np, dblocks, nd, insts, ctl, epn, epm);
    else
      hipLaunchKernelGGL(k_eval<float>, dim3(grid_of(eval_groups(np) + eval_groups(nd))),
                         dim3(kThreads), 0, s, pblocks, np, dblocks, nd, insts, ctl, epn, epm);
  }
  if (ninst > 0)
    hipLaunchKernelGGL(k_eval_finish, dim3(grid_of(ninst), kEvalOut), dim3(kThreads), 0, s, insts,
                       ctl, epn, epm, out, kEvalN, kEvalM, kEvalMaxN, kEvalMaxM, kEvalStride);
  if (np > 0) {
    if (prec == 0)
      hipLaunchKernelGGL(k_eval_ray<double>, dim3(grid_of(np)), dim3(kThreads), 0, s, pblocks,
                         insts, ctl, out, epr);
    else
      hipLaunchKernelGGL(k_eval_ray<float>, dim3(grid_of(np)), dim3(kThreads), 0, s, pblocks,
                         insts, ctl, out, epr);
  }
  if (ninst > 0)
    hipLaunchKernelGGL(k_eval_ray_finish, dim3(grid_of(ninst)), dim3(kThreads), 0, s, insts, ctl,
                       epr, out);
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

namespace pdlp {

int launch_refine(hipStream_t s, int prec, int phase, int np, const BlockDesc *pblocks, int nd,
                  const BlockDesc *dblocks, const LongRow *longs, int32_t *long_cnt,
                  double *long_part, const InstDev *insts, const Ctl *ctl, double *epn,
                  double *epm, int ninst, double *out) {
  const dim3 g(grid_of(np + nd)), b(kThreads);
  if (np + nd <= 0) return 0;
  if (phase == 1) {
    if (prec == 0)
      hipLaunchKernelGGL(k_ref_update<double>, g, b, 0, s, pblocks, np, dblocks, insts, ctl);
    else
      hipLaunchKernelGGL(k_ref_update<float>, g, b, 0, s, pblocks, np, dblocks, insts, ctl);
    hipLaunchKernelGGL(k_ref_spmv, g, b, 0, s, pblocks, np, dblocks, longs, long_cnt, long_part,
                       insts, ctl);
    hipLaunchKernelGGL(k_ref_stats, g, b, 0, s, pblocks, np, dblocks, insts, ctl, epn, epm);
    hipLaunchKernelGGL(k_eval_finish, dim3(grid_of(ninst), kRefN + kRefM), dim3(kThreads), 0, s,
                       insts, ctl, epn, epm, out, kRefN, kRefM, kRefMaxN, kRefMaxM,
                       kRefN + kRefM);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL(k_ref_setup<double>, g, b, 0, s, pblocks, np, dblocks, insts, ctl);
    else
      hipLaunchKernelGGL(k_ref_setup<float>, g, b, 0, s, pblocks, np, dblocks, insts, ctl);
  }
  return (int)hipGetLastError();
}

}  // namespace pdlp
