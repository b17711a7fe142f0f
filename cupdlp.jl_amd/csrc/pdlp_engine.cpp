// pdlp_engine.cpp — host driver of the packed restarted-PDHG engine.
#include "pdlp_engine.h"

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <cstdio>

#include "pdlp_hostsum.h"
#include "pdlp_kernels.h"
#include "pdlp_precond.h"

namespace pdlp {

double now_sec() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static inline double jl_max(double a, double b) {
  if (a != a || b != b) return a + b;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
      return -(int)(e_ == 0 ? 1 : e_);                                            \
    }                                                                             \
  } while (0)

// ---------------------------------------------------------------- pool
DevicePool::DevicePool(void *base, size_t bytes)
    : region_((char *)base), region_bytes_(bytes & ~(size_t)255) {
  free_.push_back({region_, region_bytes_});
}

DevicePool::~DevicePool() { release_all(); }

void *DevicePool::get(size_t bytes) {
  std::lock_guard<std::mutex> lk(mu_);
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  if (region_) {  // first fit in the region's free list (sorted by address)
    for (size_t i = 0; i < free_.size(); ++i) {
      if (free_[i].bytes < bytes) continue;
      void *p = free_[i].p;
      free_[i].p = (char *)free_[i].p + bytes;
      free_[i].bytes -= bytes;
      if (free_[i].bytes == 0) free_.erase(free_.begin() + i);
      used_.push_back({p, bytes});
      in_use_ += bytes;
      peak_ = std::max(peak_, in_use_);
      return p;
    }
    return nullptr;
  }
  size_t best = (size_t)-1;
  for (size_t i = 0; i < free_.size(); ++i)
    if (free_[i].bytes >= bytes && free_[i].bytes <= 2 * bytes + (1 << 20) &&
        (best == (size_t)-1 || free_[i].bytes < free_[best].bytes))
      best = i;
  if (best != (size_t)-1) {
    Blk b = free_[best];
    free_.erase(free_.begin() + best);
    used_.push_back(b);
    in_use_ += b.bytes;
    return b.p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    // release cached blocks and retry once
    for (auto &b : free_) (void)hipFree(b.p);
    free_.clear();
    (void)hipGetLastError();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  used_.push_back({p, bytes});
  in_use_ += bytes;
  peak_ = std::max(peak_, in_use_);
  return p;
}

void DevicePool::put(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  for (size_t i = 0; i < used_.size(); ++i)
    if (used_[i].p == p) {
      in_use_ -= used_[i].bytes;
      const Blk b = used_[i];
      used_.erase(used_.begin() + i);
      if (!region_) {
        free_.push_back(b);
        return;
      }
      // region: insert in address order and merge with the neighbours
      size_t k = 0;
      while (k < free_.size() && free_[k].p < b.p) ++k;
      free_.insert(free_.begin() + k, b);
      if (k + 1 < free_.size() && (char *)free_[k].p + free_[k].bytes == free_[k + 1].p) {
        free_[k].bytes += free_[k + 1].bytes;
        free_.erase(free_.begin() + k + 1);
      }
      if (k > 0 && (char *)free_[k - 1].p + free_[k - 1].bytes == free_[k].p) {
        free_[k - 1].bytes += free_[k].bytes;
        free_.erase(free_.begin() + k);
      }
      return;
    }
}

void DevicePool::release_all() {
  std::lock_guard<std::mutex> lk(mu_);
  if (region_) {  // the region is not ours to free
    free_.assign(1, Blk{region_, region_bytes_});
    used_.clear();
    in_use_ = 0;
    return;
  }
  for (auto &b : free_) (void)hipFree(b.p);
  for (auto &b : used_) (void)hipFree(b.p);
  free_.clear();
  used_.clear();
  in_use_ = 0;
}

// ---------------------------------------------------------------- engine
Engine::Engine(hipStream_t s, DevicePool *pool, const pdlp_params &prm, Profiler *prof)
    : s_(s), pool_(pool), prof_(prof) , prm_(prm) {
  prec_ = prm.precision == PDLP_PRECISION_FP32_REFINE ? 1 : 0;
  // LDS reserved per slice-kernel work-group for the occupancy cap, from the
  // current device's LDS per CU and per work-group
  int dev = 0, per_cu = 0, per_block = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
          hipSuccess &&
      hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
          hipSuccess) {
    const int po = prm.primal_occupancy > 0 ? prm.primal_occupancy : kPrimalOccupancy;
    const int dov = prm.dual_occupancy > 0 ? prm.dual_occupancy : kDualOccupancy;
    primal_lds_ = slice_lds_bytes(po, prm.slice_waves, per_cu, per_block);
    dual_lds_ = slice_lds_bytes(dov, prm.slice_waves, per_cu, per_block);
  }
}

Engine::~Engine() {
  (void)hipStreamSynchronize(s_);
  for (auto &h : inst_) {
    if (h->arena) pool_->put(h->arena);
    for (void *p : h->stage.bufs) pool_->put(p);
  }
  for (void *p : table_allocs_) pool_->put(p);
  pool_->put(d_red_);
  pool_->put(d_grow_);
  host_free(ctl_pinned_);
  host_free(out_pinned_);
  host_free(norm_pinned_);
  host_free(rout_pinned_);
  if (d_rout_) pool_->put(d_rout_);
}

// Host staging buffers of the device copies: pinned (hipHostMalloc), or
// ordinary heap memory when the engine must not allocate through HIP.
hipError_t Engine::host_alloc(void **p, size_t bytes) {
  if (!pageable_host_) return hipHostMalloc(p, bytes, 0);
  *p = malloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}

void Engine::host_free(void *p) {
  if (!p) return;
  if (pageable_host_)
    free(p);
  else
    (void)hipHostFree(p);
}

namespace {
struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  }
};
}  // namespace

namespace {
struct AfterScatterCtx {
  double *a_w, *a_o, *t_w, *t_o;
  int64_t nnz;
};
hipError_t keep_originals(hipStream_t s, void *p) {
  auto *c = (AfterScatterCtx *)p;
  if (c->nnz == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(c->a_o, c->a_w, 8 * c->nnz, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->t_o, c->t_w, 8 * c->nnz, hipMemcpyDeviceToDevice, s);
  return e;
}
}  // namespace

// Allocates the instance's arena and copies the problem into it: from the
// host (h.P) or, with src, from another instance's originals on this device.
// With device preparation the rescaling inputs stay staged for prepare(); a
// host-prepared problem is uploaded scaled and sliced here.
int Engine::transfer(HostInst &h, const DevSource *src, std::string *err) {
  Prepared &P = h.P;
  const size_t n = (size_t)P.n, m = (size_t)P.m, nnz = (size_t)P.nnz;
  const size_t ts = prec_ ? sizeof(float) : sizeof(double);
  Carver cv;
  const size_t o_aptr = cv.take(4 * (m + 1)), o_acol = cv.take(4 * nnz), o_aval = cv.take(ts * nnz);
  const size_t o_tptr = cv.take(4 * (n + 1)), o_tcol = cv.take(4 * nnz), o_tval = cv.take(ts * nnz);
  // sliced copies of A^T and A for the step kernels
  const size_t tsl = (size_t)P.ps.slots, asl = (size_t)P.ds.slots;
  const size_t o_tscol = cv.take(4 * tsl), o_tsval = cv.take(ts * tsl);
  const size_t o_ascol = cv.take(4 * asl), o_asval = cv.take(ts * asl);
  // n-vectors: c l u x0 x1 aty dx sx saty avg_x avg_g gbuf lx lg (T, except
  // the weighted sums sx, saty: FP64 on both paths)
  size_t o_n[14];
  for (int k = 0; k < 14; ++k) o_n[k] = cv.take((k == 7 || k == 8 ? 8 : ts) * n);
  // m-vectors: b y0 y1 ax0 ax1 sy sax avg_y avg_ax ly lax (T, except the
  // products ax0, ax1 and the sums sy, sax: FP64)
  size_t o_m[11];
  for (int k = 0; k < 11; ++k) o_m[k] = cv.take((k >= 3 && k <= 6 ? 8 : ts) * m);
  // double: oc ol ou D (n); ob E (m)
  size_t o_dn[4], o_dm[2];
  for (int k = 0; k < 4; ++k) o_dn[k] = cv.take(8 * n);
  for (int k = 0; k < 2; ++k) o_dm[k] = cv.take(8 * m);
  // FP32 path: the unscaled FP64 matrix in both layouts, the true problem
  // vectors and the refinement state (refinement residuals)
  size_t o_ao = 0, o_to = 0, o_ylo = 0, o_oylo = 0, o_ref[10] = {0};
  if (prec_) {
    o_ao = cv.take(8 * nnz);
    o_to = cv.take(8 * nnz);
    o_ylo = cv.take(ts * m);
    o_oylo = cv.take(8 * m);
    // xk raty rc rl ru xb (n); yk rax rb yb (m)
    for (int k = 0; k < 6; ++k) o_ref[k] = cv.take(8 * n);
    for (int k = 6; k < 10; ++k) o_ref[k] = cv.take(8 * m);
  }
  h.arena_bytes = cv.off;
  h.arena = pool_->get(cv.off);
  if (!h.arena) {
    if (err) *err = "device allocation of " + std::to_string(cv.off) + " bytes failed";
    return -2;
  }
  char *base = (char *)h.arena;
  HIPCHK(hipMemsetAsync(base, 0, cv.off, s_));
  if (src && !P.device_prep) {
    if (err) *err = "internal: staging from the device needs device preparation";
    return 8;
  }
  // host -> device, or device -> device from src's arrays
  auto put = [&](void *dst, const void *host, const void *dev, size_t bytes) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    if (src) return hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToDevice, s_);
    return hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, s_);
  };
  auto hput = [&](void *dst, const void *host, size_t bytes) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, s_);
  };
  InstDev &d = h.dev;
  d.n = (int32_t)n; d.m = (int32_t)m; d.neq = (int32_t)P.neq;
  d.a_ptr = (const int32_t *)(base + o_aptr); d.a_col = (const int32_t *)(base + o_acol);
  d.a_val = base + o_aval;
  d.t_ptr = (const int32_t *)(base + o_tptr); d.t_col = (const int32_t *)(base + o_tcol);
  d.t_val = base + o_tval;
  d.ts_col = (const int32_t *)(base + o_tscol); d.ts_val = base + o_tsval;
  d.as_col = (const int32_t *)(base + o_ascol); d.as_val = base + o_asval;
  d.c = base + o_n[0]; d.l = base + o_n[1]; d.u = base + o_n[2];
  d.x0 = base + o_n[3]; d.x1 = base + o_n[4]; d.aty = base + o_n[5]; d.dx = base + o_n[6];
  d.sx = base + o_n[7]; d.saty = base + o_n[8]; d.avg_x = base + o_n[9]; d.avg_g = base + o_n[10];
  d.gbuf = base + o_n[11]; d.lx = base + o_n[12]; d.lg = base + o_n[13];
  d.b = base + o_m[0]; d.y0 = base + o_m[1]; d.y1 = base + o_m[2]; d.ax0 = base + o_m[3];
  d.ax1 = base + o_m[4]; d.sy = base + o_m[5]; d.sax = base + o_m[6]; d.avg_y = base + o_m[7];
  d.avg_ax = base + o_m[8]; d.ly = base + o_m[9]; d.lax = base + o_m[10];
  d.oc = (const double *)(base + o_dn[0]); d.ol = (const double *)(base + o_dn[1]);
  d.ou = (const double *)(base + o_dn[2]); d.D = (const double *)(base + o_dn[3]);
  d.ob = (const double *)(base + o_dm[0]); d.E = (const double *)(base + o_dm[1]);
  d.a_orig = prec_ ? (const double *)(base + o_ao) : nullptr;
  d.t_orig = prec_ ? (const double *)(base + o_to) : nullptr;
  d.ylo = prec_ ? (void *)(base + o_ylo) : nullptr;
  d.oylo = prec_ ? (const double *)(base + o_oylo) : nullptr;
  d.xk = prec_ ? (double *)(base + o_ref[0]) : nullptr;
  d.raty = prec_ ? (double *)(base + o_ref[1]) : nullptr;
  d.rc = prec_ ? (const double *)(base + o_ref[2]) : nullptr;
  d.rl = prec_ ? (const double *)(base + o_ref[3]) : nullptr;
  d.ru = prec_ ? (const double *)(base + o_ref[4]) : nullptr;
  d.xb = prec_ ? (double *)(base + o_ref[5]) : nullptr;
  d.yk = prec_ ? (double *)(base + o_ref[6]) : nullptr;
  d.rax = prec_ ? (double *)(base + o_ref[7]) : nullptr;
  d.rb = prec_ ? (const double *)(base + o_ref[8]) : nullptr;
  d.yb = prec_ ? (double *)(base + o_ref[9]) : nullptr;
  if (prec_) {
    HIPCHK(put((void *)d.rc, P.oc.data(), src ? src->c : nullptr, 8 * n));
    HIPCHK(put((void *)d.rl, P.ol.data(), src ? src->l : nullptr, 8 * n));
    HIPCHK(put((void *)d.ru, P.ou.data(), src ? src->u : nullptr, 8 * n));
    HIPCHK(put((void *)d.rb, P.ob.data(), src ? src->b : nullptr, 8 * m));
  }
  HIPCHK(put(base + o_aptr, P.a_ptr.data(), src ? src->a_ptr : nullptr, 4 * (m + 1)));
  HIPCHK(put(base + o_tptr, P.t_ptr.data(), src ? src->t_ptr : nullptr, 4 * (n + 1)));
  HIPCHK(put(base + o_tcol, P.t_col.data(), src ? src->t_col : nullptr, 4 * nnz));
  HIPCHK(put((void *)d.oc, P.oc.data(), src ? src->c : nullptr, 8 * n));
  HIPCHK(put((void *)d.ol, P.ol.data(), src ? src->l : nullptr, 8 * n));
  HIPCHK(put((void *)d.ou, P.ou.data(), src ? src->u : nullptr, 8 * n));
  HIPCHK(put((void *)d.ob, P.ob.data(), src ? src->b : nullptr, 8 * m));

  if (P.device_prep) {
    // FP64 working copies the rescaling starts from: the arena's own arrays
    // on the FP64 path, pool buffers on the FP32 path (converted at the end)
    HostInst::Stage &sg = h.stage;
    auto scratch = [&](size_t bytes) -> double * {
      void *p = pool_->get(std::max<size_t>(bytes, 256));
      if (p) sg.bufs.push_back(p);
      return (double *)p;
    };
    if (prec_) {
      sg.t_val = scratch(8 * nnz);
      sg.c = scratch(8 * n); sg.l = scratch(8 * n); sg.u = scratch(8 * n);
      sg.b = scratch(8 * m);
    } else {
      sg.t_val = (double *)d.t_val;
      sg.c = (double *)d.c; sg.l = (double *)d.l; sg.u = (double *)d.u; sg.b = (double *)d.b;
    }
    // the slice tables (A^T first, then A): the row kernels of the
    // rescaling walk them, and they are reused to fill the sliced matrices
    const size_t nps = P.ps.sl.size(), nds = P.ds.sl.size();
    sg.tab = (SliceDesc *)scratch(sizeof(SliceDesc) * (nps + nds));
    if (!sg.t_val || !sg.c || !sg.l || !sg.u || !sg.b || !sg.tab) {
      if (err) *err = "device allocation for preconditioning failed";
      return -2;
    }
    HIPCHK(hput(sg.tab, P.ps.sl.data(), sizeof(SliceDesc) * nps));
    HIPCHK(hput(sg.tab + nps, P.ds.sl.data(), sizeof(SliceDesc) * nds));
    HIPCHK(put(sg.t_val, P.t_val.data(), src ? src->t_val : nullptr, 8 * nnz));
    HIPCHK(put(sg.c, P.oc.data(), src ? src->c : nullptr, 8 * n));
    HIPCHK(put(sg.l, P.ol.data(), src ? src->l : nullptr, 8 * n));
    HIPCHK(put(sg.u, P.ou.data(), src ? src->u : nullptr, 8 * n));
    HIPCHK(put(sg.b, P.ob.data(), src ? src->b : nullptr, 8 * m));
    P.D.assign(n, 1.0);  // the scaling starts from ones
    P.E.assign(m, 1.0);
    HIPCHK(hput((void *)d.D, P.D.data(), 8 * n));
    HIPCHK(hput((void *)d.E, P.E.data(), 8 * m));
    HIPCHK(hipStreamSynchronize(s_));
    // host copy of the values not needed any more
    std::vector<double>().swap(P.t_val);
    return 0;
  }

  // host-prepared problem: upload the scaled data
  std::vector<float> tmpf;
  auto putT = [&](void *dst, const std::vector<double> &v) -> hipError_t {
    if (v.empty()) return hipSuccess;
    if (!prec_) return hput(dst, v.data(), 8 * v.size());
    tmpf.assign(v.begin(), v.end());
    hipError_t e = hput(dst, tmpf.data(), 4 * v.size());
    if (e == hipSuccess) e = hipStreamSynchronize(s_);
    return e;
  };
  HIPCHK(hput(base + o_acol, P.a_col.data(), 4 * nnz));
  if (prec_) {  // unscaled FP64 values in both layouts
    std::vector<double> ao(nnz), to(nnz);
    for (size_t k = 0; k < nnz; ++k) to[k] = P.t_orig_host.empty() ? 0.0 : P.t_orig_host[k];
    for (size_t p = 0; p < nnz; ++p) ao[p] = to[P.csr_from_csc[p]];
    HIPCHK(hput((void *)d.t_orig, to.data(), 8 * nnz));
    HIPCHK(hput((void *)d.a_orig, ao.data(), 8 * nnz));
    HIPCHK(hipStreamSynchronize(s_));
  }
  HIPCHK(putT(base + o_aval, P.a_val));
  HIPCHK(putT(base + o_tval, P.t_val));
  HIPCHK(putT((void *)d.c, P.c));
  HIPCHK(putT((void *)d.l, P.l));
  HIPCHK(putT((void *)d.u, P.u));
  HIPCHK(putT(d.gbuf, P.c));  // gbuf = c - A^T 0
  HIPCHK(putT(d.lg, P.c));    // last restart gradient
  HIPCHK(putT((void *)d.b, P.b));
  HIPCHK(hput((void *)d.D, P.D.data(), 8 * n));
  HIPCHK(hput((void *)d.E, P.E.data(), 8 * m));
  HIPCHK(hipStreamSynchronize(s_));
  return fill_slices(h, err);
}

// Device preparation of a staged instance: rescaling (Ruiz, Pock-Chambolle)
// and the CSR scatter on the device, the FP32 conversions, the sliced
// matrices.  Returns the staged buffers to the pool.
int Engine::prepare(HostInst &h, std::string *err) {
  Prepared &P = h.P;
  if (!P.device_prep) return 0;
  const size_t n = (size_t)P.n, m = (size_t)P.m, nnz = (size_t)P.nnz;
  const size_t ts = prec_ ? sizeof(float) : sizeof(double);
  InstDev &d = h.dev;
  HostInst::Stage &sg = h.stage;
  if (!norm_pinned_) HIPCHK(host_alloc((void **)&norm_pinned_, 8 * 5 * 256));
  std::vector<void *> tmp;
  auto scratch = [&](size_t bytes) -> void * {
    void *p = pool_->get(std::max<size_t>(bytes, 256));
    if (p) tmp.push_back(p);
    return p;
  };
  double *aw = prec_ ? (double *)scratch(8 * nnz) : (double *)d.a_val;
  double *rv = (double *)scratch(8 * m), *vv = (double *)scratch(8 * n);
  int32_t *cursor = (int32_t *)scratch(4 * m);
  double *npart = (double *)scratch(8 * 5 * 256);
  auto release = [&]() {
    for (void *p : tmp) pool_->put(p);
    for (void *p : sg.bufs) pool_->put(p);
    sg = HostInst::Stage();
  };
  if (!aw || !rv || !vv || !cursor || !npart) {
    release();
    if (err) *err = "device allocation for preconditioning failed";
    return -2;
  }
  hipError_t e = m ? hipMemsetAsync(cursor, 0, 4 * m, s_) : hipSuccess;
  PrecondArgs pa;
  pa.n = (int64_t)n; pa.m = (int64_t)m; pa.nnz = (int64_t)nnz;
  pa.t_ptr = d.t_ptr; pa.t_col = d.t_col; pa.t_val = sg.t_val;
  pa.a_ptr = d.a_ptr; pa.a_col = (int32_t *)d.a_col; pa.cursor = cursor; pa.a_val = aw;
  pa.c = sg.c; pa.l = sg.l; pa.u = sg.u; pa.b = sg.b;
  pa.D = (double *)d.D; pa.E = (double *)d.E; pa.r = rv; pa.v = vv;
  pa.norm_part = npart; pa.norm_host = norm_pinned_;
  pa.ps = sg.tab; pa.nps = (int)P.ps.sl.size();
  pa.ds = sg.tab + P.ps.sl.size(); pa.nds = (int)P.ds.sl.size();
  pa.ruiz_iterations = prm_.l_inf_ruiz_iterations;
  pa.has_pc = prm_.has_pock_chambolle_alpha != 0;
  pa.pc_alpha = prm_.pock_chambolle_alpha;
  AfterScatterCtx ac{aw, (double *)d.a_orig, sg.t_val, (double *)d.t_orig, (int64_t)nnz};
  if (prec_) {
    pa.after_scatter = keep_originals;
    pa.after_scatter_ctx = &ac;
  }
  PrecondOut po;
  if (e == hipSuccess) e = precond_run(s_, pa, &po);
  if (e == hipSuccess && prec_) {
    e = to_float(s_, (int64_t)nnz, sg.t_val, (float *)d.t_val);
    if (e == hipSuccess) e = to_float(s_, (int64_t)nnz, aw, (float *)d.a_val);
    if (e == hipSuccess) e = to_float(s_, (int64_t)n, sg.c, (float *)d.c);
    if (e == hipSuccess) e = to_float(s_, (int64_t)n, sg.l, (float *)d.l);
    if (e == hipSuccess) e = to_float(s_, (int64_t)n, sg.u, (float *)d.u);
    if (e == hipSuccess) e = to_float(s_, (int64_t)m, sg.b, (float *)d.b);
  }
  if (e == hipSuccess && n) e = hipMemcpyAsync(d.gbuf, d.c, ts * n, hipMemcpyDeviceToDevice, s_);
  if (e == hipSuccess && n) e = hipMemcpyAsync(d.lg, d.c, ts * n, hipMemcpyDeviceToDevice, s_);
  if (e == hipSuccess && n) e = hipMemcpyAsync(P.D.data(), d.D, 8 * n, hipMemcpyDeviceToHost, s_);
  if (e == hipSuccess && m) e = hipMemcpyAsync(P.E.data(), d.E, 8 * m, hipMemcpyDeviceToHost, s_);
  if (e == hipSuccess) e = hipStreamSynchronize(s_);
  int rc = 0;
  if (e == hipSuccess) rc = fill_slices(h, err, sg.tab);
  release();
  if (e != hipSuccess) {
    if (err) *err = std::string("device preconditioning failed: ") + hipGetErrorString(e);
    return -(int)e;
  }
  if (rc) return rc;
  P.max_abs_scaled = po.max_abs;
  P.norm_b_scaled = po.norm_b;
  P.norm_c_scaled = po.norm_c;
  return 0;
}

// Builds the sliced copies of the scaled A^T and A from the device CSR.
// dev_tab: the instance's slice tables already on the device (A^T's, then
// A's), or null to upload them here.
int Engine::fill_slices(HostInst &h, std::string *err, const SliceDesc *dev_tab) {
  const Prepared &P = h.P;
  const InstDev &d = h.dev;
  const size_t np = P.ps.sl.size(), nd = P.ds.sl.size();
  if (np + nd == 0) return 0;
  SliceDesc *own = nullptr;
  const SliceDesc *tab = dev_tab;
  hipError_t e = hipSuccess;
  if (!tab) {
    own = (SliceDesc *)pool_->get(sizeof(SliceDesc) * (np + nd));
    if (!own) {
      if (err) *err = "device allocation of the slice tables failed";
      return -2;
    }
    if (np) e = hipMemcpyAsync(own, P.ps.sl.data(), sizeof(SliceDesc) * np, hipMemcpyHostToDevice, s_);
    if (e == hipSuccess && nd)
      e = hipMemcpyAsync(own + np, P.ds.sl.data(), sizeof(SliceDesc) * nd, hipMemcpyHostToDevice, s_);
    tab = own;
  }
  if (e == hipSuccess)
    e = (hipError_t)launch_slice_fill(s_, prec_, (int)np, tab, d.t_ptr, d.t_col, d.t_val,
                                      (int32_t *)d.ts_col, (void *)d.ts_val);
  if (e == hipSuccess)
    e = (hipError_t)launch_slice_fill(s_, prec_, (int)nd, tab + np, d.a_ptr, d.a_col, d.a_val,
                                      (int32_t *)d.as_col, (void *)d.as_val);
  if (e == hipSuccess) e = hipStreamSynchronize(s_);
  if (own) pool_->put(own);
  if (e != hipSuccess) {
    if (err) *err = std::string("building the sliced matrices failed: ") + hipGetErrorString(e);
    return -(int)e;
  }
  return 0;
}

// power_method_failure_probability — src/primal_dual_hybrid_gradient_gpu.jl:139-148
static double power_failure_probability(int64_t dim, double eps, int64_t k) {
  if (k < 2 || eps <= 0.0) return 1.0;
  const double a = 0.354 / sqrt(eps * (double)(k - 1));
  return (a < 0.824 ? a : 0.824) * sqrt((double)dim) * pow(1.0 - eps, (double)k - 0.5);
}

// SplitMix64 (the published generator of Steele, Lea and Flood) and
// Box-Muller: the power method's start vector, x_j ~ N(0, 1) for the input
// columns j = 0, 1, ... in order (the reference draws randn from
// MersenneTwister(1), which is not reproduced).
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unit_uniform(uint64_t *s) {
  return ((double)(splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// The constant step policy's step size (1 - 0.2) / sigma_max(A~), with
// sigma_max from the power method on the device
// (src/primal_dual_hybrid_gradient_gpu.jl:573-582); *passes gets the number
// of power iterations (cumulative_kkt_passes).
int Engine::constant_step(HostInst &h, double *eta, double *passes, std::string *err) {
  const Prepared &P = h.P;
  const int64_t n = P.n, m = P.m;
  const double rel = 0.2, eps = 1.0 - (1.0 - rel) * (1.0 - rel);
  int64_t iters = 0;
  while (power_failure_probability(n, eps, iters) > 0.001) ++iters;
  std::vector<double> x0(n), xd(std::max<int64_t>(n, 1));
  uint64_t st = 1;
  for (int64_t j = 0; j < n; ++j) {
    const double u1 = unit_uniform(&st), u2 = unit_uniform(&st);
    x0[j] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  for (int64_t j = 0; j < n; ++j) xd[j] = x0[P.col_perm[j]];  // device numbering
  const size_t nps = P.ps.sl.size(), nds = P.ds.sl.size();
  std::vector<void *> tmp;
  auto get = [&](size_t bytes) -> void * {
    void *p = pool_->get(std::max<size_t>(bytes, 256));
    if (p) tmp.push_back(p);
    return p;
  };
  SliceDesc *tab = (SliceDesc *)get(sizeof(SliceDesc) * (nps + nds));
  double *x = (double *)get(8 * n), *t = (double *)get(8 * m), *z = (double *)get(8 * n);
  double *part = (double *)get(8 * 512);
  hipError_t e = hipSuccess;
  if (!tab || !x || !t || !z || !part) {
    for (void *p : tmp) pool_->put(p);
    if (err) *err = "device allocation for the power method failed";
    return -2;
  }
  if (!norm_pinned_) e = host_alloc((void **)&norm_pinned_, 8 * 5 * 256);
  if (e == hipSuccess && nps)
    e = hipMemcpyAsync(tab, P.ps.sl.data(), sizeof(SliceDesc) * nps, hipMemcpyHostToDevice, s_);
  if (e == hipSuccess && nds)
    e = hipMemcpyAsync(tab + nps, P.ds.sl.data(), sizeof(SliceDesc) * nds, hipMemcpyHostToDevice, s_);
  if (e == hipSuccess && n) e = hipMemcpyAsync(x, xd.data(), 8 * n, hipMemcpyHostToDevice, s_);
  double sv = 0.0;
  if (e == hipSuccess) {
    PowerArgs a;
    a.n = n; a.m = m;
    a.ps = tab; a.nps = (int)nps; a.ds = tab + nps; a.nds = (int)nds;
    a.t_ptr = h.dev.t_ptr; a.t_col = h.dev.t_col; a.t_val = (const double *)h.dev.t_val;
    a.a_ptr = h.dev.a_ptr; a.a_col = h.dev.a_col; a.a_val = (const double *)h.dev.a_val;
    a.x = x; a.t = t; a.z = z; a.part = part; a.part_host = norm_pinned_;
    a.iterations = iters;
    e = power_method(s_, a, &sv);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s_);
  for (void *p : tmp) pool_->put(p);
  if (e != hipSuccess) {
    if (err) *err = std::string("power method failed: ") + hipGetErrorString(e);
    return -(int)e;
  }
  *eta = (1.0 - rel) / sv;
  *passes = (double)iters;
  return 0;
}

int Engine::stage(std::unique_ptr<HostInst> h, std::string *err, const DevSource *src) {
  if (prm_.step_size_policy == PDLP_STEP_CONSTANT && prec_) {
    if (err) *err = "ConstantStepsizeParams needs PDLP_PRECISION_FP64";
    return 4;
  }
  int rc = transfer(*h, src, err);
  if (rc) {
    (void)hipStreamSynchronize(s_);
    for (void *p : h->stage.bufs) pool_->put(p);
    if (h->arena) pool_->put(h->arena);
    return rc;
  }
  h->staged = true;
  Ctl c;
  memset(&c, 0, sizeof c);
  inst_.push_back(std::move(h));
  ctl_.push_back(c);
  tables_dirty_ = true;
  return 0;
}

int Engine::add(std::unique_ptr<HostInst> h, std::string *err) {
  int rc = stage(std::move(h), err);
  if (rc) return rc;
  rc = activate(inst_.size() - 1, err);
  if (rc) {  // the engine does not keep an instance it could not start
    (void)hipStreamSynchronize(s_);
    HostInst &b = *inst_.back();
    for (void *p : b.stage.bufs) pool_->put(p);
    if (b.arena) pool_->put(b.arena);
    inst_.pop_back();
    ctl_.pop_back();
  }
  return rc;
}

int Engine::activate_staged(std::string *err) {
  for (size_t q = 0; q < inst_.size(); ++q) {
    if (!inst_[q]->staged || inst_[q]->done) continue;
    int rc = activate(q, err);
    if (rc) return rc;
  }
  return 0;
}

// Prepares a staged instance on the device and sets its initial state.
int Engine::activate(size_t q, std::string *err) {
  HostInst *h = inst_[q].get();
  int rc = prepare(*h, err);
  if (rc) return rc;
  h->staged = false;
  const bool constant = prm_.step_size_policy == PDLP_STEP_CONSTANT;
  Ctl c;
  memset(&c, 0, sizeof c);
  const Prepared &P = h->P;
  // 1 / norm(A, Inf) with the elementwise norm (src/primal_dual_hybrid_gradient_gpu.jl:569-571);
  // a matrix with no entries (m = 0) uses 1.0, where the reference fails.
  double eta0 = P.max_abs_scaled > 0.0 ? 1.0 / P.max_abs_scaled : 1.0;
  double passes0 = 0.5;
  if (constant) {
    rc = constant_step(*h, &eta0, &passes0, err);
    if (rc) return rc;
  }
  c.eta = eta0;
  c.eta_entry = eta0;
  double omega = prm_.primal_importance;
  if (prm_.scale_invariant_initial_primal_weight) {
    // select_initial_primal_weight — src/saddle_point_gpu.jl:767-785
    const double nb = P.device_prep ? P.norm_b_scaled : dd_norm2(P.b);
    const double nc = P.device_prep ? P.norm_c_scaled : dd_norm2(P.c);
    const double rhs = sqrt(1.0) * nb, obj = sqrt(1.0) * nc;
    if (obj > 0.0 && rhs > 0.0) omega = prm_.primal_importance * (obj / rhs);
  }
  c.omega = omega;
  c.kkt_passes = passes0;
  c.aty_valid = 1;
  h->qc_cur = P.qc;
  if (prec_) h->tol_cur = prm_.refine_inner_tolerance;
  h->needs_eval = true;
  h->eval_iter = 1;
  h->start_time = now_sec();
  ctl_[q] = c;
  tables_dirty_ = true;
  return 0;
}

static void partition_rows(const std::vector<int32_t> &ptr, int64_t R, int inst,
                           std::vector<BlockDesc> &blocks, std::vector<LongRow> &longs,
                           int *nslots) {
  if (R == 0) {  // one empty block, so every instance takes part in each launch
    blocks.push_back({inst, 0, 0, 0});
    return;
  }
  int64_t r = 0;
  while (r < R) {
    const int64_t len = ptr[r + 1] - ptr[r];
    if (len > kBlockNnz) {
      const int nch = (int)((len + kLongChunk - 1) / kLongChunk);
      const int lid = (int)longs.size();
      longs.push_back({*nslots, nch});
      *nslots += nch;
      for (int c = 0; c < nch; ++c) blocks.push_back({inst, (int32_t)r, -(c + 1), lid});
      ++r;
      continue;
    }
    const int64_t start = r;
    int64_t nz = 0;
    while (r < R && r - start < kBlockRows) {
      const int64_t l = ptr[r + 1] - ptr[r];
      if (l > kBlockNnz || nz + l > kBlockNnz) break;
      nz += l;
      ++r;
    }
    blocks.push_back({inst, (int32_t)start, (int32_t)(r - start), 0});
  }
}

// Copies the host control blocks to the device through the pinned staging
// buffer.  The stream is drained first: an earlier asynchronous copy out of
// the same buffer may not have run yet, and overwriting the buffer before it
// runs would hand the earlier kernels the later contents.
hipError_t Engine::push_ctl() {
  const size_t ninst = inst_.size();
  if (ninst == 0) return hipSuccess;
  hipError_t e = hipStreamSynchronize(s_);
  if (e != hipSuccess) return e;
  memcpy(ctl_pinned_, ctl_.data(), sizeof(Ctl) * ninst);
  return hipMemcpyAsync(d_ctl_, ctl_pinned_, sizeof(Ctl) * ninst, hipMemcpyHostToDevice, s_);
}

int Engine::rebuild_tables(std::string *err) {
  HIPCHK(hipStreamSynchronize(s_));
  for (void *p : table_allocs_) pool_->put(p);
  table_allocs_.clear();
  std::vector<BlockDesc> pb, db;
  std::vector<LongRow> longs;
  int nslots = 0;
  // Finished instances get no blocks: launches then cost only what the
  // running instances need.  An instance's own blocks do not depend on where
  // it sits in the tables, so dropping others changes none of its sums.
  for (size_t q = 0; q < inst_.size(); ++q) {
    HostInst &h = *inst_[q];
    h.dev.pb0 = (int32_t)pb.size();
    if (!h.done && !h.staged) partition_rows(h.P.t_ptr, h.P.n, (int)q, pb, longs, &nslots);
    h.dev.pb1 = (int32_t)pb.size();
  }
  for (size_t q = 0; q < inst_.size(); ++q) {
    HostInst &h = *inst_[q];
    h.dev.db0 = (int32_t)db.size();
    if (!h.done && !h.staged) partition_rows(h.P.a_ptr, h.P.m, (int)q, db, longs, &nslots);
    h.dev.db1 = (int32_t)db.size();
  }
  // slice tables of the step kernels: normal slices, then the chunk slices
  // of long rows (with LongRow entries after the row blocks' ones); the
  // chunk slices' partials follow the normal slices' in the partial arrays
  std::vector<SliceDesc> ps, ds, pls, dls;
  auto add_slices = [&](const SliceTable &st, int q, std::vector<SliceDesc> &out,
                        std::vector<SliceDesc> &lout) {
    const int lid0 = (int)longs.size();
    for (int32_t nch : st.nchunks) {
      longs.push_back({nslots, nch});
      nslots += nch;
    }
    for (SliceDesc sd : st.sl) {
      sd.inst = q;
      if (sd.nrows == 0) {
        sd.lid += lid0;
        lout.push_back(sd);
      } else {
        out.push_back(sd);
      }
    }
  };
  for (size_t q = 0; q < inst_.size(); ++q) {
    HostInst &h = *inst_[q];
    h.dev.ps0 = (int32_t)ps.size();
    h.dev.pl0 = (int32_t)pls.size();
    h.dev.ds0 = (int32_t)ds.size();
    h.dev.dl0 = (int32_t)dls.size();
    if (!h.done && !h.staged) {
      add_slices(h.P.ps, (int)q, ps, pls);
      add_slices(h.P.ds, (int)q, ds, dls);
    }
    h.dev.ps1 = (int32_t)ps.size();
    h.dev.pl1 = (int32_t)pls.size();
    h.dev.ds1 = (int32_t)ds.size();
    h.dev.dl1 = (int32_t)dls.size();
  }
  for (auto &h : inst_) {
    h->dev.pl0 += (int32_t)ps.size();
    h->dev.pl1 += (int32_t)ps.size();
    h->dev.dl0 += (int32_t)ds.size();
    h->dev.dl1 += (int32_t)ds.size();
  }
  // partial indices: normal slices in instance order, then the chunk slices
  for (size_t k = 0; k < ps.size(); ++k) ps[k].part = (int32_t)k;
  for (size_t k = 0; k < ds.size(); ++k) ds[k].part = (int32_t)k;
  for (size_t k = 0; k < pls.size(); ++k) pls[k].part = (int32_t)(ps.size() + k);
  for (size_t k = 0; k < dls.size(); ++k) dls[k].part = (int32_t)(ds.size() + k);
  const size_t nparts_p = ps.size() + pls.size(), nparts_d = ds.size() + dls.size();
  npl_ = (int)pls.size();
  ndl_ = (int)dls.size();
  ps.insert(ps.end(), pls.begin(), pls.end());  // one upload: normal slices, then chunks
  ds.insert(ds.end(), dls.begin(), dls.end());
  np_ = (int)pb.size();
  nd_ = (int)db.size();
  nps_ = (int)ps.size() - npl_;
  nds_ = (int)ds.size() - ndl_;
  nlong_ = (int)longs.size();
  nslots_ = nslots;
  const size_t ninst = inst_.size();
  auto alloc = [&](size_t bytes) -> void * {
    void *p = pool_->get(std::max<size_t>(bytes, 256));
    if (p) table_allocs_.push_back(p);
    return p;
  };
  d_pslices_ = (SliceDesc *)alloc(sizeof(SliceDesc) * ps.size());
  d_dslices_ = (SliceDesc *)alloc(sizeof(SliceDesc) * ds.size());
  d_pblocks_ = (BlockDesc *)alloc(sizeof(BlockDesc) * pb.size());
  d_dblocks_ = (BlockDesc *)alloc(sizeof(BlockDesc) * db.size());
  d_longs_ = (LongRow *)alloc(sizeof(LongRow) * longs.size());
  d_long_cnt_ = (int32_t *)alloc(sizeof(int32_t) * longs.size());
  d_long_part_ = (double *)alloc(2 * sizeof(double) * nslots);
  d_insts_ = (InstDev *)alloc(sizeof(InstDev) * ninst);
  d_ctl_ = (Ctl *)alloc(sizeof(Ctl) * ninst);
  // compensated partial sums: (s, e) pairs
  d_ppart_ = (double *)alloc(sizeof(double) * 2 * nparts_p);
  d_dpart_ = (double *)alloc(sizeof(double) * 4 * nparts_d);
  d_epn_ = (double *)alloc(sizeof(double) * 2 * kEvalN * pb.size());
  d_epm_ = (double *)alloc(sizeof(double) * 2 * kEvalM * db.size());
  d_epr_ = (double *)alloc(sizeof(double) * 2 * pb.size());
  d_out_ = (double *)alloc(sizeof(double) * kEvalStride * ninst);
  if (!d_pslices_ || !d_dslices_ || !d_pblocks_ || !d_dblocks_ || !d_longs_ || !d_long_cnt_ || !d_long_part_ || !d_insts_ ||
      !d_ctl_ || !d_ppart_ || !d_dpart_ || !d_epn_ || !d_epm_ || !d_epr_ || !d_out_) {
    if (err) *err = "device allocation of block tables failed";
    return -2;
  }
  if (!ps.empty())
    HIPCHK(hipMemcpyAsync(d_pslices_, ps.data(), sizeof(SliceDesc) * ps.size(), hipMemcpyHostToDevice, s_));
  if (!ds.empty())
    HIPCHK(hipMemcpyAsync(d_dslices_, ds.data(), sizeof(SliceDesc) * ds.size(), hipMemcpyHostToDevice, s_));
  if (!pb.empty())
    HIPCHK(hipMemcpyAsync(d_pblocks_, pb.data(), sizeof(BlockDesc) * pb.size(), hipMemcpyHostToDevice, s_));
  if (!db.empty())
    HIPCHK(hipMemcpyAsync(d_dblocks_, db.data(), sizeof(BlockDesc) * db.size(), hipMemcpyHostToDevice, s_));
  if (!longs.empty()) {
    HIPCHK(hipMemcpyAsync(d_longs_, longs.data(), sizeof(LongRow) * longs.size(), hipMemcpyHostToDevice, s_));
    HIPCHK(hipMemsetAsync(d_long_cnt_, 0, sizeof(int32_t) * longs.size(), s_));
  }
  std::vector<InstDev> devs(ninst);
  for (size_t q = 0; q < ninst; ++q) devs[q] = inst_[q]->dev;
  HIPCHK(hipMemcpyAsync(d_insts_, devs.data(), sizeof(InstDev) * ninst, hipMemcpyHostToDevice, s_));
  host_free(ctl_pinned_);
  ctl_pinned_ = nullptr;
  HIPCHK(host_alloc((void **)&ctl_pinned_, sizeof(Ctl) * std::max<size_t>(ninst, 1)));
  if (out_cap_ < ninst) {
    host_free(out_pinned_);
    out_pinned_ = nullptr;
    HIPCHK(host_alloc((void **)&out_pinned_, sizeof(double) * kEvalStride * std::max<size_t>(ninst, 1)));
    out_cap_ = ninst;
  }
  HIPCHK(push_ctl());
  HIPCHK(hipStreamSynchronize(s_));
  tables_dirty_ = false;
  return 0;
}

int Engine::ensure_step_table(int64_t need, std::string *err) {
  if (need < tab_len_) return 0;
  int64_t len = std::max<int64_t>(65536, tab_len_);
  while (len <= need) len *= 2;
  std::vector<double> red(len), grow(len);
  // (1 - (k+1)^-r) and (1 + (k+1)^-g) for attempt number k —
  // src/primal_dual_hybrid_gradient_gpu.jl:405-409
  for (int64_t k = 0; k < len; ++k) {
    const double k1 = (double)(k + 1);
    red[k] = 1.0 - 1.0 / pow(k1, prm_.reduction_exponent);
    grow[k] = 1.0 + 1.0 / pow(k1, prm_.growth_exponent);
  }
  HIPCHK(hipStreamSynchronize(s_));
  pool_->put(d_red_);
  pool_->put(d_grow_);
  d_red_ = (double *)pool_->get(sizeof(double) * len);
  d_grow_ = (double *)pool_->get(sizeof(double) * len);
  if (!d_red_ || !d_grow_) {
    if (err) *err = "device allocation of the step table failed";
    return -2;
  }
  HIPCHK(hipMemcpy(d_red_, red.data(), sizeof(double) * len, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_grow_, grow.data(), sizeof(double) * len, hipMemcpyHostToDevice));
  tab_len_ = len;
  return 0;
}

// Enqueues `count` step attempts of every running instance and the copy of
// the control blocks back to the host; complete_attempts() waits for them.
int Engine::launch_attempts(int64_t count, std::string *err) {
  int64_t max_att = 0;
  for (size_t q = 0; q < inst_.size(); ++q) max_att = std::max(max_att, ctl_[q].attempts);
  int rc = ensure_step_table(max_att + count + 2, err);
  if (rc) return rc;
  const bool prof = prof_ && prof_->enabled;
  std::vector<double> &pbytes = pend_pbytes_, &dbytes = pend_dbytes_;
  std::vector<int64_t> &remaining = pend_remaining_;
  pbytes.clear();
  dbytes.clear();
  remaining.clear();
  // Algorithmic bytes per launch of each instance: the matrix with its
  // indices, the row pointers, the gathered vector read once, and nine fused
  // vector streams per row (primal: x c l u in, x' out, the two weighted sums
  // read and written; dual: y b Ax in, y' Ax' out, two sums read and written):
  //   N (s_A + 4) + 4 (rows + 1) + s_v (gathered) + 9 s_v rows,
  // with s_A = s_v = 8 (FP64) or 4 (FP32).  Traffic the engine adds beyond
  // this (A^T y and dx stored, FP64 sums on the FP32 path) is not counted.
  const bool timed = prof && prof_->timed;
  if (prof) {
    const size_t need = timed ? 4 * (size_t)count : 0;
    while (prof_->events.size() < need) {
      hipEvent_t ev;
      HIPCHK(hipEventCreate(&ev));
      prof_->events.push_back(ev);
    }
    for (size_t q = 0; q < inst_.size(); ++q) {
      const Prepared &P = inst_[q]->P;
      const double N = (double)P.nnz, n = (double)P.n, m = (double)P.m;
      const double sv = prec_ ? 4.0 : 8.0;
      pbytes.push_back(N * (sv + 4) + 4 * (n + 1) + sv * m + 9 * sv * n);
      dbytes.push_back(N * (sv + 4) + 4 * (m + 1) + sv * n + 9 * sv * m);
      remaining.push_back(inst_[q]->done ? 0 : ctl_[q].target - ctl_[q].accepted);
    }
  }
  const double t0 = now_sec();
  for (int64_t a = 0; a < count; ++a) {
    hipEvent_t *ev = timed ? &prof_->events[4 * a] : nullptr;
    int e = launch_primal(s_, prec_, nps_, d_pslices_, npl_, d_pslices_ + nps_, d_longs_,
                          d_long_cnt_, d_long_part_, d_insts_, d_ctl_, d_ppart_,
                          prm_.slice_unroll, prm_.slice_waves, primal_lds_,
                          ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    if (!e)
      e = launch_dual(s_, prec_, nds_, d_dslices_, ndl_, d_dslices_ + nds_, d_longs_,
                      d_long_cnt_, d_long_part_, d_insts_, d_ctl_, d_dpart_, prm_.slice_unroll,
                      prm_.slice_waves, dual_lds_, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
    if (!e)
      e = launch_step(s_, (int)inst_.size(), d_insts_, d_ctl_, d_ppart_, d_dpart_, d_red_, d_grow_,
                      tab_len_, prm_.reduction_exponent, prm_.growth_exponent,
                      prm_.step_size_policy == PDLP_STEP_CONSTANT ? 1 : 0,
                      prm_.debug_numerical_error_attempt);
    if (e) {
      if (err) *err = std::string("kernel launch failed: ") + hipGetErrorString((hipError_t)e);
      return -e;
    }
  }
  const size_t ninst = inst_.size();
  HIPCHK(hipMemcpyAsync(ctl_pinned_, d_ctl_, sizeof(Ctl) * ninst, hipMemcpyDeviceToHost, s_));
  pend_count_ = count;
  pend_t0_ = t0;
  pend_prof_ = prof;
  return 0;
}

int Engine::complete_attempts(std::string *err) {
  const int64_t count = pend_count_;
  if (count <= 0) return 0;
  pend_count_ = 0;
  const bool prof = pend_prof_;
  const std::vector<double> &pbytes = pend_pbytes_, &dbytes = pend_dbytes_;
  const std::vector<int64_t> &remaining = pend_remaining_;
  const size_t ninst = inst_.size();
  HIPCHK(hipStreamSynchronize(s_));
  memcpy(ctl_.data(), ctl_pinned_, sizeof(Ctl) * ninst);
  const double dt = now_sec() - pend_t0_;
  if (prof) {
    for (int64_t a = 0; a < count; ++a) {
      float ms_p = 0.f, ms_d = 0.f;
      if (prof_->timed) {
        HIPCHK(hipEventElapsedTime(&ms_p, prof_->events[4 * a + 0], prof_->events[4 * a + 1]));
        HIPCHK(hipEventElapsedTime(&ms_d, prof_->events[4 * a + 2], prof_->events[4 * a + 3]));
      }
      double bp = 0.0, bd = 0.0;
      for (size_t q = 0; q < ninst; ++q)
        if (a < remaining[q]) {
          bp += pbytes[q];
          bd += dbytes[q];
        }
      prof_->acc.primal_launches += 1;
      prof_->acc.dual_launches += 1;
      prof_->acc.primal_ms += ms_p;
      prof_->acc.dual_ms += ms_d;
      prof_->acc.primal_bytes += bp;
      prof_->acc.dual_bytes += bd;
    }
  }
  for (size_t q = 0; q < ninst; ++q) {
    HostInst &h = *inst_[q];
    if (h.done) continue;
    h.basic_time += dt;
    Ctl &C = ctl_[q];
    if (C.numerical_error) {
      h.eval_iter = h.iter_start + (C.accepted - h.acc_at_eval) + 1;
      h.needs_eval = true;
    } else if (C.accepted >= C.target) {
      h.needs_eval = true;
    }
  }
  return 0;
}

// gap_norm >= 0 replaces |primal objective| + |dual objective| in the gap test
// (correction solves of the refinement path).
static int check_termination(const pdlp_termination_criteria &tc, const QpCache &qc,
                             const pdlp_iteration_stats &st, double gap_norm = -1.0) {
  const pdlp_convergence_information &ci = st.convergence_information;
  const double abs_obj = gap_norm >= 0.0 ? gap_norm
                                         : fabs(ci.primal_objective) + fabs(ci.dual_objective);
  const double gap = fabs(ci.primal_objective - ci.dual_objective);
  double pe, pb, de, db;
  if (tc.optimality_norm == PDLP_L_INF) {
    pe = ci.l_inf_primal_residual; pb = qc.l_inf_b; de = ci.l_inf_dual_residual; db = qc.l_inf_c;
  } else {
    pe = ci.l2_primal_residual; pb = qc.l2_b; de = ci.l2_dual_residual; db = qc.l2_c;
  }
  const double ea = tc.eps_optimal_absolute, er = tc.eps_optimal_relative;
  if (de < ea + er * db && pe < ea + er * pb && gap < ea + er * abs_obj)
    return PDLP_TERMINATION_REASON_OPTIMAL;
  const pdlp_infeasibility_information &ii = st.infeasibility_information;
  if (!(ii.dual_ray_objective <= 0.0) &&
      ii.max_dual_ray_infeasibility / ii.dual_ray_objective <= tc.eps_primal_infeasible)
    return PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE;
  if (!(ii.primal_ray_linear_objective >= 0.0) &&
      ii.max_primal_ray_infeasibility / (-ii.primal_ray_linear_objective) <= tc.eps_dual_infeasible)
    return PDLP_TERMINATION_REASON_DUAL_INFEASIBLE;
  if ((int64_t)st.iteration_number >= (int64_t)tc.iteration_limit)
    return PDLP_TERMINATION_REASON_ITERATION_LIMIT;
  if (st.cumulative_kkt_matrix_passes >= tc.kkt_matrix_pass_limit)
    return PDLP_TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT;
  if (st.cumulative_time_sec >= tc.time_sec_limit) return PDLP_TERMINATION_REASON_TIME_LIMIT;
  return 0;
}

// Iteration cap of the first FP32 solve of FP32 + refinement.
constexpr int64_t kFirstSolveCap = 100000;

static int64_t next_eval_iteration(int64_t it, int64_t freq, int64_t limit) {
  int64_t j = it + 1;
  while (!((j - 1) % freq == 0 || j == limit + 1 || j <= 10)) ++j;
  return j;
}

void Engine::host_eval_logic(size_t q, const double *o) {
  HostInst &h = *inst_[q];
  Ctl &C = ctl_[q];
  const Prepared &P = h.P;
  const double *N = o, *M = o + kEvalN;
  pdlp_termination_criteria tc = prm_.termination_criteria;
  const double tol = h.tol_cur > 0 ? h.tol_cur : h.tolerance_override;
  if (tol > 0) {
    tc.eps_optimal_absolute = tol;
    tc.eps_optimal_relative = tol;
  }
  if (h.iteration_limit_override >= 0)
    tc.iteration_limit = (int32_t)std::min<int64_t>(tc.iteration_limit, h.iteration_limit_override);
  if (h.time_limit_override >= 0) tc.time_sec_limit = std::min(tc.time_sec_limit, h.time_limit_override);
  // FP32 + refinement: a correction solve counts its iterations from 1, so
  // its limit is what is left of the caller's after the earlier rounds (the
  // next evaluation point and the limit test both use it)
  if (prec_)
    tc.iteration_limit =
        (int32_t)std::max<int64_t>((int64_t)tc.iteration_limit - h.iters_total, 0);
  const QpCache &qc = h.qc_cur;
  const double c0 = P.c0;
  C.kkt_passes += 2.0;
  const double eps_ratio = tc.eps_optimal_absolute / tc.eps_optimal_relative;
  pdlp_iteration_stats st;
  memset(&st, 0, sizeof st);
  st.iteration_number = (int32_t)(h.eval_iter - 1);
  st.cumulative_kkt_matrix_passes = C.kkt_passes;
  st.cumulative_time_sec = now_sec() - h.start_time;
  // compute_convergence_information — src/iteration_stats_utils_gpu.jl:251-309
  pdlp_convergence_information &ci = st.convergence_information;
  ci.candidate_type = PDLP_POINT_TYPE_AVERAGE_ITERATE;
  ci.primal_objective = c0 + N[2];
  if (prec_) ci.primal_objective -= M[18];  // -ylo' max(Ax - b, 0) of the correction LP
  ci.l_inf_primal_residual = jl_max(M[1], N[1]);
  ci.l2_primal_residual = sqrt(M[0] + N[0]);
  ci.relative_l_inf_primal_residual = ci.l_inf_primal_residual / (eps_ratio + qc.l_inf_b);
  ci.relative_l2_primal_residual = ci.l2_primal_residual / (eps_ratio + qc.l2_b);
  ci.l_inf_primal_variable = N[3];
  ci.l2_primal_variable = sqrt(N[4]);
  const double dri = jl_max(M[3], N[6]);
  ci.dual_objective = (M[4] + c0) + N[7];
  ci.l_inf_dual_residual = dri;
  ci.l2_dual_residual = sqrt(M[2] + N[5]);
  ci.relative_l_inf_dual_residual = ci.l_inf_dual_residual / (eps_ratio + qc.l_inf_c);
  ci.relative_l2_dual_residual = ci.l2_dual_residual / (eps_ratio + qc.l2_c);
  ci.l_inf_dual_variable = M[5];
  ci.l2_dual_variable = sqrt(M[6]);
  ci.corrected_dual_objective = dri == 0.0 ? ci.dual_objective : -INFINITY;
  {
    const double gap = fabs(ci.primal_objective - ci.dual_objective);
    const double abs_obj = fabs(ci.primal_objective) + fabs(ci.dual_objective);
    ci.relative_optimality_gap = gap / (eps_ratio + abs_obj);
  }
  // compute_infeasibility_information — src/iteration_stats_utils_gpu.jl:314-372
  pdlp_infeasibility_information &ii = st.infeasibility_information;
  ii.candidate_type = PDLP_POINT_TYPE_AVERAGE_ITERATE;
  {
    const double s = N[3];
    const double bmax = jl_max(M[7], N[11]);
    ii.max_primal_ray_infeasibility = s != 0.0 ? bmax / s : bmax;
    ii.primal_ray_linear_objective = s != 0.0 ? o[kEvalOut] : N[2];
    const double dri2 = jl_max(M[3], N[8]);
    const double dobj2 = (M[4] + c0) + N[9];
    const double scaling = jl_max(M[5], N[10]);
    if (scaling != 0.0) {
      ii.max_dual_ray_infeasibility = dri2 / scaling;
      ii.dual_ray_objective = dobj2 / scaling;
    }
  }
  st.step_size = C.eta_entry;
  st.primal_weight = C.omega;
  st.time_spent_doing_basic_algorithm = h.basic_time;
  st.total_attempts = C.attempts;

  int reason = check_termination(tc, qc, st, h.gap_norm_cur);
  if (C.numerical_error && reason == 0) reason = PDLP_TERMINATION_REASON_NUMERICAL_ERROR;
  // a correction solve that has not met its tolerance within its cap is
  // refined anyway (the outer rounds are bounded by refine_max_outer)
  if (reason == 0 && prec_ && h.inner_cap > 0 && (int64_t)st.iteration_number >= h.inner_cap) {
    reason = PDLP_TERMINATION_REASON_ITERATION_LIMIT;
    ++h.capped_rounds;
  }
  // the first FP32 solve is capped too (an inner tolerance FP32 data cannot
  // reach would otherwise never end it): the refinement then counts as
  // stalled and the caller's FP64 fallback takes over
  if (reason == 0 && prec_ && h.ref_round == 0 &&
      (int64_t)st.iteration_number >=
          std::max<int64_t>(kFirstSolveCap, prm_.refine_inner_iteration_limit)) {
    reason = PDLP_TERMINATION_REASON_ITERATION_LIMIT;
    ++h.capped_rounds;
  }
  if (h.eval_iter < 10 && (reason == PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE ||
                           reason == PDLP_TERMINATION_REASON_DUAL_INFEASIBLE))
    reason = 0;
  // the flush applied the pending sums and A^T y
  C.pending = 0;
  C.aty_valid = 1;
  C.restart_action = 0;
  if (reason != 0) {
    h.last_stats = st;
    C.active = 0;
    const bool final_reason = reason == PDLP_TERMINATION_REASON_TIME_LIMIT ||
                              reason == PDLP_TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT ||
                              ((reason == PDLP_TERMINATION_REASON_PRIMAL_INFEASIBLE ||
                                reason == PDLP_TERMINATION_REASON_DUAL_INFEASIBLE) &&
                               h.ref_round == 0);
    if (prec_ && !final_reason) {
      // the FP32 solve of the current (correction) LP is done: refine
      h.refine_now = true;
      h.iters_total += h.eval_iter - 1;
      h.attempts_total += C.attempts;
      return;
    }
    h.done = true;
    h.reason = reason;
    if (prec_) {
      h.iters_total += h.eval_iter - 1;
      h.attempts_total += C.attempts;
    }
    if (h.cb && prm_.record_iteration_stats && !prec_) h.cb(&st, h.user);
    return;
  }
  // run_restart_scheme — src/saddle_point_gpu.jl:479-622 (scaled space)
  const pdlp_restart_params &rp = prm_.restart_params;
  const double omega = C.omega;
  const double pnp = 1.0 / C.eta_entry * omega;
  const double dnp = 1.0 / C.eta_entry / omega;
  auto kkt = [&](int c) {
    const double l2p = sqrt(M[8 + 3 * c] + N[12 + 4 * c]);
    double pobj = c0 + N[13 + 4 * c];
    if (prec_) pobj -= M[19 + c];
    const double l2d = sqrt(M[9 + 3 * c] + N[14 + 4 * c]);
    const double dobj = (M[10 + 3 * c] + c0) + N[15 + 4 * c];
    const double gap = fabs(pobj - dobj);
    return sqrt(omega * (l2p * l2p) + 1.0 / omega * (l2d * l2d) + gap * gap);
  };
  int32_t restart_used = PDLP_RESTART_CHOICE_NO_RESTART;
  if (C.count > 0) {
    const int64_t len = C.count;
    bool do_restart = (double)len >= rp.artificial_restart_threshold * (double)(h.eval_iter - 1);
    bool reset_to_average = false;
    double cand = NAN;
    if (rp.restart_scheme != PDLP_NO_RESTARTS) {
      const double cur = kkt(0), avg = kkt(1);
      reset_to_average = rp.restart_to_current_metric == PDLP_KKT_GREEDY ? (cur >= avg) : true;
      cand = reset_to_average ? avg : cur;
    }
    if (!do_restart) {
      if (rp.restart_scheme == PDLP_ADAPTIVE_KKT) {
        const double ratio = cand / kkt(2);
        if (ratio < rp.necessary_reduction_for_restart) {
          if (ratio < rp.sufficient_reduction_for_restart)
            do_restart = true;
          else if (ratio > h.last_trial)
            do_restart = true;
        }
        h.last_trial = ratio;
      } else if (rp.restart_scheme == PDLP_FIXED_FREQUENCY &&
                 rp.restart_frequency_if_fixed <= len) {
        do_restart = true;
      }
    }
    if (do_restart) {
      C.restart_action = reset_to_average ? 2 : 1;
      h.pdist = sqrt(pnp) * sqrt(N[24]) / sqrt(omega);
      h.ddist = sqrt(dnp) * sqrt(M[17]) * sqrt(omega);
      C.count = 0;
      C.wsum = 0.0;
      restart_used = reset_to_average ? PDLP_RESTART_CHOICE_RESTART_TO_AVERAGE
                                      : PDLP_RESTART_CHOICE_WEIGHTED_AVERAGE_RESET;
    }
  }
  st.restart_used = restart_used;
  if (restart_used != PDLP_RESTART_CHOICE_NO_RESTART) {
    // compute_new_primal_weight — src/saddle_point_gpu.jl:627-654
    const double eps = 2.220446049250313e-16;
    if (h.pdist > eps && h.ddist > eps) {
      const double s = rp.primal_weight_update_smoothing;
      const double est = h.ddist / h.pdist;
      const double lw = s * log(est) + (1.0 - s) * log(C.omega);
      C.omega = exp(lw);
    }
  }
  h.last_stats = st;
  if (h.cb && prm_.record_iteration_stats && !prec_) h.cb(&st, h.user);
  const int64_t next = next_eval_iteration(h.eval_iter, prm_.termination_evaluation_frequency,
                                           tc.iteration_limit);
  h.iter_start = h.eval_iter;
  h.acc_at_eval = C.accepted;
  C.target = C.accepted + (next - h.eval_iter);
  h.eval_iter = next;
  C.active = 1;
  h.needs_eval = false;
}

int Engine::evaluate(std::string *err) {
  const size_t ninst = inst_.size();
  int neval = 0;
  for (size_t q = 0; q < ninst; ++q) {
    HostInst &h = *inst_[q];
    Ctl &C = ctl_[q];
    const bool ev = !h.done && h.needs_eval;
    C.eval_flag = ev ? 1 : 0;
    C.restart_action = 0;
    if (ev) {
      C.active = 0;
      C.use_current = (C.numerical_error || C.count == 0) ? 1 : 0;
      ++neval;
    }
  }
  if (neval == 0) return 0;
  HIPCHK(push_ctl());
  int e = launch_primal_flush(s_, prec_, np_, d_pblocks_, d_longs_, d_long_cnt_, d_long_part_,
                              d_insts_, d_ctl_);
  if (!e) e = launch_dual_flush(s_, prec_, nd_, d_dblocks_, d_insts_, d_ctl_);
  if (!e)
    e = launch_eval(s_, prec_, np_, d_pblocks_, nd_, d_dblocks_, d_insts_, d_ctl_, d_epn_, d_epm_,
                    d_epr_, (int)ninst, d_out_);
  if (e) {
    if (err) *err = std::string("evaluation launch failed: ") + hipGetErrorString((hipError_t)e);
    return -e;
  }
  HIPCHK(hipMemcpyAsync(out_pinned_, d_out_, sizeof(double) * kEvalStride * ninst,
                        hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  for (size_t q = 0; q < ninst; ++q)
    if (ctl_[q].eval_flag) host_eval_logic(q, out_pinned_ + q * kEvalStride);
  HIPCHK(push_ctl());
  e = launch_restart(s_, prec_, np_, d_pblocks_, nd_, d_dblocks_, d_insts_, d_ctl_);
  if (e) {
    if (err) *err = std::string("restart launch failed: ") + hipGetErrorString((hipError_t)e);
    return -e;
  }
  if (prec_) {
    int rc = refine(err);
    if (rc) return rc;
  }
  for (size_t q = 0; q < ninst; ++q) {
    if (ctl_[q].eval_flag && inst_[q]->done) {
      int rc = finish_instance(q, err);
      if (rc) return rc;
    }
  }
  return 0;
}

int Engine::finish_instance(size_t q, std::string *err) {
  HostInst &h = *inst_[q];
  const Ctl &C = ctl_[q];
  pdlp_result *r = h.res;
  if (!r) return 0;
  const size_t n = (size_t)h.P.n, m = (size_t)h.P.m;
  const size_t ts = prec_ ? 4 : 8;
  std::vector<char> bx(ts * std::max<size_t>(n, 1)), by(ts * std::max<size_t>(m, 1));
  auto fetch = [&](const void *src, std::vector<char> &dst, size_t cnt) -> hipError_t {
    if (cnt == 0) return hipSuccess;
    return hipMemcpyAsync(dst.data(), src, ts * cnt, hipMemcpyDeviceToHost, s_);
  };
  auto val = [&](const std::vector<char> &b, size_t i) -> double {
    return prec_ ? (double)((const float *)b.data())[i] : ((const double *)b.data())[i];
  };
  // device entry j is the caller's entry col_perm[j] (rows: row_perm)
  const int32_t *cp = h.P.col_perm.data(), *rp = h.P.row_perm.data();
  if (h.refined) {  // the refined FP64 point, or the best one if the last round made it worse
    const bool best = h.last_err > h.best_err;
    std::vector<double> hx(std::max<size_t>(n, 1)), hy(std::max<size_t>(m, 1));
    if (n) HIPCHK(hipMemcpyAsync(hx.data(), best ? h.dev.xb : h.dev.xk, 8 * n, hipMemcpyDeviceToHost, s_));
    if (m) HIPCHK(hipMemcpyAsync(hy.data(), best ? h.dev.yb : h.dev.yk, 8 * m, hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
    for (size_t j = 0; j < n; ++j) r->x[cp[j]] = hx[j];
    for (size_t i = 0; i < m; ++i) r->y[rp[i]] = hy[i];
    if (r->x_current && n) memcpy(r->x_current, r->x, 8 * n);
    if (r->y_current && m) memcpy(r->y_current, r->y, 8 * m);
    r->termination_reason = h.reason;
    r->solve_time_sec = now_sec() - h.t_submit;
    r->iteration_count = (int32_t)h.iters_total;
    r->final_stats = best ? h.best_stats : h.last_stats;
    r->total_attempts = h.attempts_total;
    r->refinement_rounds = h.ref_round;
    return 0;
  }
  HIPCHK(fetch(h.dev.avg_x, bx, n));
  HIPCHK(fetch(h.dev.avg_y, by, m));
  HIPCHK(hipStreamSynchronize(s_));
  for (size_t j = 0; j < n; ++j) r->x[cp[j]] = val(bx, j) / h.P.D[j];
  for (size_t i = 0; i < m; ++i) r->y[rp[i]] = val(by, i) / h.P.E[i];
  if (r->x_current || r->y_current) {
    HIPCHK(fetch(C.cur ? h.dev.x1 : h.dev.x0, bx, n));
    HIPCHK(fetch(C.cur ? h.dev.y1 : h.dev.y0, by, m));
    HIPCHK(hipStreamSynchronize(s_));
    if (r->x_current)
      for (size_t j = 0; j < n; ++j) r->x_current[cp[j]] = val(bx, j) / h.P.D[j];
    if (r->y_current)
      for (size_t i = 0; i < m; ++i) r->y_current[rp[i]] = val(by, i) / h.P.E[i];
  }
  r->termination_reason = h.reason;
  r->solve_time_sec = now_sec() - h.t_submit;
  r->iteration_count = (int32_t)(prec_ ? h.iters_total : h.eval_iter - 1);
  r->final_stats = h.last_stats;
  r->total_attempts = prec_ ? h.attempts_total : C.attempts;
  r->refinement_rounds = h.ref_round;
  return 0;
}

// Refinement of the instances whose FP32 solve just finished: update the FP64
// point, evaluate it on the true problem in FP64, then either finish or set
// up the next correction LP.
int Engine::refine(std::string *err) {
  const size_t ninst = inst_.size();
  int nref = 0;
  for (size_t q = 0; q < ninst; ++q) {
    HostInst &h = *inst_[q];
    Ctl &C = ctl_[q];
    C.ref_flag = h.refine_now ? 1 : 0;  // eval_flag is already set for these
    if (h.refine_now) {
      C.ref_dp = h.dp;
      C.ref_dd = h.dd;
      ++nref;
    }
  }
  if (!nref) return 0;
  if (rout_cap_ < ninst) {
    HIPCHK(hipStreamSynchronize(s_));
    if (d_rout_) pool_->put(d_rout_);
    host_free(rout_pinned_);
    d_rout_ = nullptr;
    rout_pinned_ = nullptr;
    rout_cap_ = 0;
    d_rout_ = (double *)pool_->get(sizeof(double) * kRefOutStride * ninst);
    if (!d_rout_) {
      if (err) *err = "device allocation of refinement statistics failed";
      return -2;
    }
    HIPCHK(host_alloc((void **)&rout_pinned_, sizeof(double) * kRefOutStride * ninst));
    rout_cap_ = ninst;
  }
  HIPCHK(push_ctl());
  int e = launch_refine(s_, prec_, 1, np_, d_pblocks_, nd_, d_dblocks_, d_longs_, d_long_cnt_,
                        d_long_part_, d_insts_, d_ctl_, d_epn_, d_epm_, (int)ninst, d_rout_);
  if (e) {
    if (err) *err = std::string("refinement launch failed: ") + hipGetErrorString((hipError_t)e);
    return -e;
  }
  HIPCHK(hipMemcpyAsync(rout_pinned_, d_rout_, sizeof(double) * kRefOutStride * ninst,
                        hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  for (size_t q = 0; q < ninst; ++q)
    if (inst_[q]->refine_now) refine_logic(q, rout_pinned_ + q * kRefOutStride);
  HIPCHK(push_ctl());
  e = launch_refine(s_, prec_, 2, np_, d_pblocks_, nd_, d_dblocks_, d_longs_, d_long_cnt_,
                    d_long_part_, d_insts_, d_ctl_, d_epn_, d_epm_, (int)ninst, d_rout_);
  if (e) {
    if (err) *err = std::string("refinement launch failed: ") + hipGetErrorString((hipError_t)e);
    return -e;
  }
  for (size_t q = 0; q < ninst; ++q) {
    HostInst &h = *inst_[q];
    if (!h.refine_now) continue;
    h.refine_now = false;
    ctl_[q].ref_flag = 0;
  }
  return 0;
}

// Outer test on the refined point (optimality_criteria_met, src/termination.jl:142-170,
// in the original space) and the scale factors of the next correction LP.
void Engine::refine_logic(size_t q, const double *o) {
  HostInst &h = *inst_[q];
  Ctl &C = ctl_[q];
  const Prepared &P = h.P;
  const double *N = o, *M = o + kRefN;
  h.refined = true;
  pdlp_termination_criteria tc = prm_.termination_criteria;
  if (h.tolerance_override > 0) {
    tc.eps_optimal_absolute = h.tolerance_override;
    tc.eps_optimal_relative = h.tolerance_override;
  }
  if (h.iteration_limit_override >= 0)
    tc.iteration_limit = (int32_t)std::min<int64_t>(tc.iteration_limit, h.iteration_limit_override);
  if (h.time_limit_override >= 0) tc.time_sec_limit = std::min(tc.time_sec_limit, h.time_limit_override);
  const double eps_ratio = tc.eps_optimal_absolute / tc.eps_optimal_relative;
  const double c0 = P.c0;
  pdlp_iteration_stats st;
  memset(&st, 0, sizeof st);
  st.iteration_number = (int32_t)h.iters_total;
  st.cumulative_kkt_matrix_passes = C.kkt_passes;
  st.cumulative_time_sec = now_sec() - h.start_time;
  pdlp_convergence_information &ci = st.convergence_information;
  ci.candidate_type = PDLP_POINT_TYPE_AVERAGE_ITERATE;
  ci.primal_objective = c0 + N[2];
  ci.l_inf_primal_residual = jl_max(M[1], N[1]);
  ci.l2_primal_residual = sqrt(M[0] + N[0]);
  ci.relative_l_inf_primal_residual = ci.l_inf_primal_residual / (eps_ratio + P.qc.l_inf_b);
  ci.relative_l2_primal_residual = ci.l2_primal_residual / (eps_ratio + P.qc.l2_b);
  ci.l_inf_primal_variable = N[3];
  ci.l2_primal_variable = sqrt(N[4]);
  const double dri = jl_max(M[3], N[6]);
  ci.dual_objective = (M[4] + c0) + N[7];
  ci.l_inf_dual_residual = dri;
  ci.l2_dual_residual = sqrt(M[2] + N[5]);
  ci.relative_l_inf_dual_residual = ci.l_inf_dual_residual / (eps_ratio + P.qc.l_inf_c);
  ci.relative_l2_dual_residual = ci.l2_dual_residual / (eps_ratio + P.qc.l2_c);
  ci.l_inf_dual_variable = M[5];
  ci.l2_dual_variable = sqrt(M[6]);
  ci.corrected_dual_objective = dri == 0.0 ? ci.dual_objective : -INFINITY;
  const double gap = fabs(ci.primal_objective - ci.dual_objective);
  const double abs_obj = fabs(ci.primal_objective) + fabs(ci.dual_objective);
  ci.relative_optimality_gap = gap / (eps_ratio + abs_obj);
  st.infeasibility_information.candidate_type = PDLP_POINT_TYPE_AVERAGE_ITERATE;
  st.step_size = C.eta_entry;
  st.primal_weight = C.omega;
  st.time_spent_doing_basic_algorithm = h.basic_time;
  st.total_attempts = h.attempts_total;
  h.last_stats = st;
  double pe, pb, de, db;
  if (tc.optimality_norm == PDLP_L_INF) {
    pe = ci.l_inf_primal_residual; pb = P.qc.l_inf_b; de = ci.l_inf_dual_residual; db = P.qc.l_inf_c;
  } else {
    pe = ci.l2_primal_residual; pb = P.qc.l2_b; de = ci.l2_dual_residual; db = P.qc.l2_c;
  }
  const double ea = tc.eps_optimal_absolute, er = tc.eps_optimal_relative;
  // the point's error as a multiple of the tolerance (optimal iff < 1)
  const double err = jl_max(jl_max(pe / (ea + er * pb), de / (ea + er * db)),
                            gap / (ea + er * abs_obj));
  const double prev_err = h.last_err;
  h.last_err = err;
  // whether the solve that produced this point stopped at its iteration cap
  const bool capped_now = h.capped_rounds > h.capped_seen;
  h.capped_seen = h.capped_rounds;
  const bool improved = err < h.best_err;
  if (improved) {
    h.best_err = err;
    h.best_stats = st;
  }
  int reason = 0;
  if (de < ea + er * db && pe < ea + er * pb && gap < ea + er * abs_obj)
    reason = PDLP_TERMINATION_REASON_OPTIMAL;
  else if (st.cumulative_time_sec >= tc.time_sec_limit)
    reason = PDLP_TERMINATION_REASON_TIME_LIMIT;
  else if (h.iters_total >= (int64_t)tc.iteration_limit)
    reason = PDLP_TERMINATION_REASON_ITERATION_LIMIT;
  else if (h.ref_round >= prm_.refine_max_outer ||
           (capped_now && !(h.ref_round > 0 && prm_.refine_continue_progress > 0 &&
                            err * (double)prm_.refine_continue_progress <= prev_err))) {
    // refinement stalled (the round limit, the first FP32 solve stopped by
    // its cap, or a correction solve stopped by its cap without cutting the
    // error refine_continue_progress times): the caller may solve the LP
    // again in FP64.  A capped correction solve that did cut the error that
    // much moved the point closer to optimal, and the next round starts
    // from there.
    reason = PDLP_TERMINATION_REASON_ITERATION_LIMIT;
    h.stalled = true;
  }
  if (reason) {
    h.done = true;
    h.reason = reason;
    C.ref_flag = 0;
    if (h.cb && prm_.record_iteration_stats) h.cb(&st, h.user);
    return;
  }
  // Scale factors: the inverse largest violations, grown by at most
  // refine_max_scale per round.  The correction LP's own tests are relative
  // to its violations (not to the full reduced costs / residuals, whose
  // legitimate parts -- multipliers of active bounds, slacks of inactive
  // rows -- say nothing about how far (xk, yk) is from optimal), and its
  // initial primal weight balances the two scaled violations.
  const double alpha = prm_.refine_max_scale > 1.0 ? prm_.refine_max_scale : 1.0;
  const double ep = jl_max(M[1], N[1]), edl = jl_max(N[6], M[3]);
  const double dp = ep > 0.0 ? std::min(1.0 / ep, alpha * h.dp) : alpha * h.dp;
  const double ddn = edl > 0.0 ? std::min(1.0 / edl, alpha * h.dd) : alpha * h.dd;
  const double dp_prev = h.dp, dd_prev = h.dd;
  h.dp = dp;
  h.dd = ddn;
  if (h.ref_round == 0)
    h.inner_cap = prm_.refine_inner_iteration_limit > 0
                      ? (int64_t)prm_.refine_inner_iteration_limit
                      : std::max<int64_t>(10000, 20 * h.iters_total);
  h.ref_round += 1;
  h.qc_cur.l_inf_b = dp * ep;
  h.qc_cur.l2_b = dp * sqrt(M[0] + N[0]);
  h.qc_cur.l_inf_c = ddn * edl;
  h.qc_cur.l2_c = ddn * sqrt(N[5] + M[2]);
  // The correction's gap is its objective error in units scaled by dP dD:
  // test it against the current gap of (xk, yk) in those units, not against
  // the correction's own objective values (small, so that test would ask for
  // far more accuracy than the refined point can use).
  h.gap_norm_cur = dp * ddn * gap;
  const double nc = ddn * sqrt(N[8] + M[8]), nbn = dp * sqrt(M[7]);
  double omega0 = prm_.primal_importance;
  if (prm_.scale_invariant_initial_primal_weight && nc > 0.0 && nbn > 0.0)
    omega0 = prm_.primal_importance * (nc / nbn);
  // refine_warm_start bit 1: the previous solve's final primal weight
  // instead, carried to the correction's scaling (a weight scales as dD / dP)
  if ((prm_.refine_warm_start & 1) && C.omega > 0.0)
    omega0 = C.omega * (ddn / dp) / (dd_prev / dp_prev);
  C.ref_flag = 2;
  C.ref_keep = improved ? 1 : 0;
  C.ref_dp = dp;
  C.ref_dd = ddn;
  if (h.cb && prm_.record_iteration_stats) h.cb(&st, h.user);
  reset_inner(q, omega0);
  // A correction solved to relative tolerance t leaves about t times the
  // current violations, so a point at err times the outer tolerance needs t
  // near 1/err: ask for ten times less, never tighter than
  // refine_inner_tolerance (nor looser than 0.1).
  if (std::isfinite(err))
    h.tol_cur = std::max(prm_.refine_inner_tolerance, std::min(0.1, 0.1 / err));
}

// Fresh PDHG state for the next correction LP (the device side is reset by
// the setup kernel).
void Engine::reset_inner(size_t q, double omega0) {
  HostInst &h = *inst_[q];
  Ctl &C = ctl_[q];
  // refine_warm_start bit 2: the previous solve's final step size (the
  // correction LP has the same scaled matrix)
  const double eta0 = (prm_.refine_warm_start & 2) && C.eta_entry > 0.0 ? C.eta_entry
                      : h.P.max_abs_scaled > 0.0 ? 1.0 / h.P.max_abs_scaled
                                                 : 1.0;
  const double kkt = C.kkt_passes;
  const double dp = C.ref_dp, ddn = C.ref_dd;
  const int32_t flag = C.ref_flag, keep = C.ref_keep;
  memset(&C, 0, sizeof C);
  C.eta = eta0;
  C.eta_entry = eta0;
  C.omega = omega0;
  C.kkt_passes = kkt;
  C.aty_valid = 1;
  C.ref_flag = flag;
  C.ref_keep = keep;
  C.ref_dp = dp;
  C.ref_dd = ddn;
  h.eval_iter = 1;
  h.iter_start = 1;
  h.acc_at_eval = 0;
  h.needs_eval = true;
  h.last_trial = 1.0;
  h.pdist = h.ddist = 0.0;
  h.tol_cur = prm_.refine_inner_tolerance;
}

int Engine::begin_round(std::string *err) {
  int rc0 = activate_staged(err);
  if (rc0) return rc0;
  // Rebuild the block tables once finished instances hold a quarter of them.
  int64_t dead = 0;
  for (auto &h : inst_)
    if (h->done) dead += (h->dev.pb1 - h->dev.pb0) + (h->dev.db1 - h->dev.db0);
  if (4 * dead > (int64_t)np_ + nd_) tables_dirty_ = true;
  if (tables_dirty_) {
    int rc = rebuild_tables(err);
    if (rc) return rc;
  }
  bool any_eval = false;
  for (auto &h : inst_)
    if (!h->done && h->needs_eval) any_eval = true;
  if (any_eval) {
    int rc = evaluate(err);
    if (rc) return rc;
  }
  int64_t R = 0;
  bool all_done = true;
  for (size_t q = 0; q < inst_.size(); ++q) {
    if (inst_[q]->done) continue;
    all_done = false;
    R = std::max<int64_t>(R, ctl_[q].target - ctl_[q].accepted);
  }
  if (all_done) return 1;
  if (R <= 0) return 0;
  R = std::min<int64_t>(R, std::max<int32_t>(prm_.termination_evaluation_frequency, 10));
  int rc = launch_attempts(R, err);
  return rc < 0 ? rc : 0;
}

int Engine::end_round(std::string *err) {
  int rc = complete_attempts(err);
  return rc < 0 ? rc : 0;
}

int Engine::step_round(std::string *err) {
  int r = begin_round(err);
  if (r != 0) return r;
  return end_round(err);
}

bool Engine::round_ready() const {
  return pend_count_ <= 0 || hipStreamQuery(s_) == hipSuccess;
}

bool Engine::all_done() const {
  for (auto &h : inst_)
    if (!h->done) return false;
  return true;
}

int Engine::run(std::string *err) {
  for (;;) {
    int r = step_round(err);
    if (r < 0) return r;
    if (r == 1) return 0;
  }
}

int Engine::cancel(size_t q, std::string *err) {
  HostInst &h = *inst_[q];
  if (h.done) return 0;
  if (h.staged) {  // not prepared yet: no device block; release the staging buffers
    HIPCHK(hipStreamSynchronize(s_));
    for (void *p : h.stage.bufs) pool_->put(p);
    h.stage = HostInst::Stage();
    h.staged = false;
    h.done = true;
    h.needs_eval = false;
    tables_dirty_ = true;
    return 0;
  }
  h.done = true;
  h.needs_eval = false;
  ctl_[q].active = 0;
  // Only this instance's block is copied: instances staged since the last
  // table rebuild have no device block yet (the rebuild at the next round
  // copies them all), and an active instance's block exists.
  HIPCHK(hipStreamSynchronize(s_));
  HIPCHK(hipMemcpyAsync(d_ctl_ + q, &ctl_[q], sizeof(Ctl), hipMemcpyHostToDevice, s_));
  HIPCHK(hipStreamSynchronize(s_));
  return 0;
}

void Engine::remove_done() {
  std::vector<std::unique_ptr<HostInst>> keep;
  std::vector<Ctl> kc;
  for (size_t q = 0; q < inst_.size(); ++q) {
    if (inst_[q]->done) {
      if (inst_[q]->arena) pool_->put(inst_[q]->arena);
      inst_[q]->arena = nullptr;
      continue;
    }
    keep.push_back(std::move(inst_[q]));
    kc.push_back(ctl_[q]);
  }
  if (keep.size() != inst_.size()) tables_dirty_ = true;
  inst_ = std::move(keep);
  ctl_ = std::move(kc);
}

}  // namespace pdlp
