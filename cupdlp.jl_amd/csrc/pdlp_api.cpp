// pdlp_api.cpp — the C ABI of libpdlp_hip.so (declared in include/pdlp_hip.h).
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pdlp_engine.h"
#include "pdlp_hip.h"
#include "pdlp_noalloc.h"
#include "pdlp_prep.h"

using namespace pdlp;

struct pdlp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // FP64 solves of stalled refinements (created on first use)
  std::vector<hipStream_t> extra; // streams of a batch's further groups (created on first use)
  DevicePool pool;
  Profiler prof;
  std::string err;
  bool prof_fallback_timed = false;  // pdlp_set_profiling(ctx, 2)
  std::map<std::string, pdlp_tuning> tune_cache;  // pdlp_tune's choices per LP shape
  int scratch_rc = 0;             // set-up of the no-allocation region (pdlp_ctx_create)
  std::string scratch_err;
};

static std::mutex g_err_mu;
static std::string g_err;

static void set_global_error(const std::string &e) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = e;
}

extern "C" {

int pdlp_abi_version(void) { return PDLP_ABI_VERSION; }

int32_t pdlp_abi_layout(int64_t *out, int32_t max) {
  const int64_t sizes[] = {(int64_t)sizeof(pdlp_lp),
                           (int64_t)sizeof(pdlp_termination_criteria),
                           (int64_t)sizeof(pdlp_restart_params),
                           (int64_t)sizeof(pdlp_params),
                           (int64_t)sizeof(pdlp_convergence_information),
                           (int64_t)sizeof(pdlp_infeasibility_information),
                           (int64_t)sizeof(pdlp_iteration_stats),
                           (int64_t)sizeof(pdlp_result)};
  const int32_t cnt = (int32_t)(sizeof(sizes) / sizeof(sizes[0]));
  for (int32_t i = 0; i < cnt && i < max; ++i) out[i] = sizes[i];
  return cnt < max ? cnt : max;
}

void pdlp_default_params(pdlp_params *p) {
  memset(p, 0, sizeof(*p));
  // cuPDLP.Optimizer() — src/MOI_wrapper.jl:58-78
  p->l_inf_ruiz_iterations = 10;
  p->l2_norm_rescaling = 0;
  p->has_pock_chambolle_alpha = 1;
  p->pock_chambolle_alpha = 1.0;
  p->primal_importance = 1.0;
  p->scale_invariant_initial_primal_weight = 1;
  p->verbosity = 2;
  p->record_iteration_stats = 1;
  p->termination_evaluation_frequency = 64;
  p->step_size_policy = PDLP_STEP_ADAPTIVE;
  p->reduction_exponent = 0.3;
  p->growth_exponent = 0.6;
  // construct_termination_criteria() — src/termination.jl:79-99
  p->termination_criteria.optimality_norm = PDLP_L2;
  p->termination_criteria.eps_optimal_absolute = 1.0e-6;
  p->termination_criteria.eps_optimal_relative = 1.0e-6;
  p->termination_criteria.eps_primal_infeasible = 1.0e-8;
  p->termination_criteria.eps_dual_infeasible = 1.0e-8;
  p->termination_criteria.time_sec_limit = INFINITY;
  p->termination_criteria.iteration_limit = 2147483647;
  p->termination_criteria.kkt_matrix_pass_limit = INFINITY;
  p->restart_params.restart_scheme = PDLP_ADAPTIVE_KKT;
  p->restart_params.restart_to_current_metric = PDLP_KKT_GREEDY;
  p->restart_params.restart_frequency_if_fixed = 1000;
  p->restart_params.artificial_restart_threshold = 0.36;
  p->restart_params.sufficient_reduction_for_restart = 0.2;
  p->restart_params.necessary_reduction_for_restart = 0.8;
  p->restart_params.primal_weight_update_smoothing = 0.5;
  p->precision = PDLP_PRECISION_FP64;
  p->refine_max_outer = 12;
  p->refine_inner_tolerance = 1.0e-4;
  p->refine_max_scale = 1.0e6;
  p->device_preconditioning = 1;
  p->refine_inner_iteration_limit = 0;
  p->refine_warm_start = 3;
  p->refine_continue_progress = 0;
}

int pdlp_ctx_create(int device, pdlp_ctx **out) {
  if (!out) return 1;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    set_global_error(std::string("no HIP device available: ") + hipGetErrorString(e));
    return -1;
  }
  if (device < 0 || device >= ndev) {
    set_global_error("device index out of range");
    return 1;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    set_global_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
    return -(int)e;
  }
  auto *c = new pdlp_ctx();
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_global_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    delete c;
    return -(int)e;
  }
  // the device address of this device's no-allocation region (once per
  // device, the device current); a failure is reported by pdlp_solve_noalloc
  c->scratch_rc = scratch_setup(device, &c->scratch_err);
  *out = c;
  return 0;
}

void pdlp_ctx_destroy(pdlp_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t ev : ctx->prof.events) (void)hipEventDestroy(ev);
  ctx->pool.release_all();
  (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) {
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream2);
  }
  for (hipStream_t s : ctx->extra) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  delete ctx;
}

const char *pdlp_last_error(const pdlp_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

int pdlp_set_profiling(pdlp_ctx *ctx, int32_t enable) {
  if (!ctx) return 1;
  ctx->prof.enabled = enable != 0;
  ctx->prof_fallback_timed = enable >= 2;
  return 0;
}

int pdlp_get_profile(pdlp_ctx *ctx, pdlp_kernel_profile *out) {
  if (!ctx || !out) return 1;
  *out = ctx->prof.acc;
  memset(&ctx->prof.acc, 0, sizeof ctx->prof.acc);
  return 0;
}

}  // extern "C"

static int check_params(const pdlp_params *p, std::string *err) {
  const pdlp_restart_params &rp = p->restart_params;
  // construct_restart_parameters asserts — src/saddle_point_gpu.jl:397-403
  if (!(rp.restart_frequency_if_fixed > 1) ||
      !(0.0 < rp.artificial_restart_threshold && rp.artificial_restart_threshold <= 1.0) ||
      !(0.0 < rp.sufficient_reduction_for_restart &&
        rp.sufficient_reduction_for_restart <= rp.necessary_reduction_for_restart &&
        rp.necessary_reduction_for_restart <= 1.0) ||
      !(0.0 <= rp.primal_weight_update_smoothing && rp.primal_weight_update_smoothing <= 1.0)) {
    *err = "invalid restart parameters";
    return 2;
  }
  if (p->termination_evaluation_frequency <= 0) {
    *err = "termination_evaluation_frequency must be positive";
    return 2;
  }
  // src/primal_dual_hybrid_gradient_gpu.jl:483-485
  if (!(p->primal_importance > 0) || !isfinite(p->primal_importance)) {
    *err = "primal_importance must be positive and finite";
    return 3;
  }
  if (p->has_pock_chambolle_alpha && !(p->pock_chambolle_alpha >= 0 && p->pock_chambolle_alpha <= 2)) {
    *err = "pock_chambolle_alpha must be in [0, 2]";
    return 2;
  }
  if (p->precision != PDLP_PRECISION_FP64 && p->precision != PDLP_PRECISION_FP32_REFINE) {
    *err = "precision must be PDLP_PRECISION_FP64 or PDLP_PRECISION_FP32_REFINE";
    return 2;
  }
  if (p->precision == PDLP_PRECISION_FP32_REFINE &&
      (p->refine_max_outer < 0 || !(p->refine_inner_tolerance > 0) ||
       !(p->refine_max_scale >= 1.0) || p->refine_inner_iteration_limit < 0 ||
       p->refine_continue_progress < 0)) {
    *err = "invalid refinement parameters";
    return 2;
  }
  if (p->lane_budget != 0 && (p->lane_budget < 16 || p->lane_budget > 1024)) {
    *err = "lane_budget must be 0 or in [16, 1024]";
    return 2;
  }
  if ((p->slice_unroll != 0 && p->slice_unroll != 4 && p->slice_unroll != 8 &&
       p->slice_unroll != 16) ||
      (p->slice_waves != 0 && p->slice_waves != 1 && p->slice_waves != 2 && p->slice_waves != 4)) {
    *err = "slice_unroll must be 0, 4, 8 or 16 and slice_waves 0, 1, 2 or 4";
    return 2;
  }
  if (p->primal_occupancy < 0 || p->primal_occupancy > 16 || p->dual_occupancy < 0 ||
      p->dual_occupancy > 16) {
    *err = "primal_occupancy and dual_occupancy must be in [0, 16]";
    return 2;
  }
  return 0;
}

namespace pdlp {
int prepare_instance(const pdlp_lp *lp, const pdlp_params *prm, HostInst *h, std::string *err);
}  // namespace pdlp

int pdlp::prepare_instance(const pdlp_lp *lp, const pdlp_params *prm, HostInst *h,
                           std::string *err) {
  // Rescaling runs on the device unless disabled or the optional L2 pass is
  // requested (that pass is host-only).
  const bool dev = prm->device_preconditioning && !prm->l2_norm_rescaling;
  int rc = load_problem(lp, &h->P, err, !dev, prm->lane_budget);
  if (rc) return rc;
  h->P.device_prep = dev;
  if (!dev) {
    if (prm->precision == PDLP_PRECISION_FP32_REFINE) h->P.t_orig_host = h->P.t_val;
    rescale_host(prm, &h->P);
  }
  return 0;
}

// A host copy of an LP: FP32 + refinement with host rescaling keeps one per
// instance, for the FP64 solve of an instance whose refinement stalls.
struct LpCopy {
  pdlp_lp lp{};
  std::vector<int64_t> colptr, rowval;
  std::vector<double> nzval, c, lb, ub, b;
  explicit LpCopy(const pdlp_lp &in) {
    lp = in;
    const int64_t n = in.n, m = in.m, nnz = in.nnz;
    colptr.assign(in.colptr, in.colptr + n + 1);
    rowval.assign(in.rowval, in.rowval + nnz);
    nzval.assign(in.nzval, in.nzval + nnz);
    c.assign(in.c, in.c + n);
    lb.assign(in.lb, in.lb + n);
    ub.assign(in.ub, in.ub + n);
    b.assign(in.b, in.b + m);
    lp.colptr = colptr.data();
    lp.rowval = rowval.data();
    lp.nzval = nzval.data();
    lp.c = c.data();
    lp.lb = lb.data();
    lp.ub = ub.data();
    lp.b = b.data();
  }
};

// The refined point a stalled FP32 solve ended with, kept while the LP is
// solved again in FP64.  Should that solve stop at the caller's iteration,
// time or KKT-pass limit with a larger relative KKT error than the refined
// point (a tight limit leaves it few iterations), the refined point and its
// statistics are returned instead; the reason and counts stay the whole
// solve's.
struct RefinedPoint {
  std::vector<double> x, y, xc, yc;
  pdlp_iteration_stats stats{};

  void save(const pdlp_result &r, int64_t n, int64_t m) {
    x.assign(r.x, r.x + n);
    y.assign(r.y, r.y + m);
    if (r.x_current) xc.assign(r.x_current, r.x_current + n);
    if (r.y_current) yc.assign(r.y_current, r.y_current + m);
    stats = r.final_stats;
  }

  // max of the relative primal residual, dual residual and gap in the
  // termination norm: each is the absolute figure over (eps_abs / eps_rel +
  // its norm), so the max orders points as the optimality test does
  static double rel_kkt(const pdlp_iteration_stats &s, int32_t norm) {
    const pdlp_convergence_information &c = s.convergence_information;
    const bool linf = norm == PDLP_L_INF;
    const double p = linf ? c.relative_l_inf_primal_residual : c.relative_l2_primal_residual;
    const double d = linf ? c.relative_l_inf_dual_residual : c.relative_l2_dual_residual;
    const double g = c.relative_optimality_gap;
    if (isnan(p) || isnan(d) || isnan(g)) return INFINITY;
    return fmax(fmax(p, d), g);
  }

  void restore_if_better(pdlp_result *r, int32_t norm) const {
    const int32_t why = r->termination_reason;
    if (why != PDLP_TERMINATION_REASON_ITERATION_LIMIT &&
        why != PDLP_TERMINATION_REASON_TIME_LIMIT &&
        why != PDLP_TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT)
      return;
    if (!(rel_kkt(stats, norm) < rel_kkt(r->final_stats, norm))) return;
    if (!x.empty()) memcpy(r->x, x.data(), 8 * x.size());
    if (!y.empty()) memcpy(r->y, y.data(), 8 * y.size());
    if (r->x_current && !xc.empty()) memcpy(r->x_current, xc.data(), 8 * xc.size());
    if (r->y_current && !yc.empty()) memcpy(r->y_current, yc.data(), 8 * yc.size());
    r->final_stats = stats;
  }
};

// Stream k of a context, created on first use: 0 is the context's stream,
// 1 the stream of the FP64 solves of stalled refinements, 2g and 2g + 1 the
// same pair for group g of a batch.  Default priority, non-blocking.
static int ctx_stream(pdlp_ctx *ctx, int k, hipStream_t *out, std::string *err) {
  hipStream_t *slot;
  if (k == 0) {
    slot = &ctx->stream;
  } else if (k == 1) {
    slot = &ctx->stream2;
  } else {
    if ((int)ctx->extra.size() < k - 1) ctx->extra.resize(k - 1, nullptr);
    slot = &ctx->extra[k - 2];
  }
  if (!*slot) {
    hipError_t e = hipStreamCreateWithFlags(slot, hipStreamNonBlocking);
    if (e != hipSuccess) {
      *slot = nullptr;
      *err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
      return -(int)e;
    }
  }
  *out = *slot;
  return 0;
}

// Stages in `fb` the plain FP64 solve of an FP32 instance whose refinement
// stalled.  With device rescaling it starts from the FP32 instance's unscaled
// FP64 copies on the device and its host layout (the same numbering and
// slices), so no host data of the LP is needed; with host rescaling, from a
// host copy of the LP.  g carries the result pointer and limits.
static int stage_fp64(Engine *fb, const HostInst &h32, const LpCopy *lp, const pdlp_params &p64,
                      std::unique_ptr<HostInst> g, std::string *err) {
  if (lp) {
    int rc = prepare_instance(&lp->lp, &p64, g.get(), err);
    if (rc) return rc;
    return fb->stage(std::move(g), err);
  }
  g->P = h32.P;
  DevSource src;
  src.t_ptr = h32.dev.t_ptr;
  src.t_col = h32.dev.t_col;
  src.a_ptr = h32.dev.a_ptr;
  src.t_val = h32.dev.t_orig;
  // the true problem's vectors (oc, ol, ou, ob hold the correction LP's)
  src.c = h32.dev.rc;
  src.l = h32.dev.rl;
  src.u = h32.dev.ru;
  src.b = h32.dev.rb;
  return fb->stage(std::move(g), err, &src);
}

// A plain FP64 solve of an FP32 instance started while its refinement still
// runs (refine_early_fallback), into a result of its own.  It becomes the
// instance's result only if that refinement then stalls and the early start
// cannot have changed it; otherwise it is cancelled.
struct EarlyFp64 {
  int32_t k = -1;      // index in the group's FP64 engine (-1: none)
  int32_t round = 0;   // the correction round it started in
  bool committed = false;
  double commit_sec = 0.0;  // time from the call's start to the stall
  pdlp_result res{};
  std::vector<double> x, y;
};

// One group of a batch: its instances packed into one engine's launches on
// its own stream, with its own FP64 engine for stalled refinements.
struct BatchGroup {
  std::unique_ptr<Engine> main;
  std::vector<int32_t> idx;  // the caller's index of each engine instance
  std::vector<std::unique_ptr<LpCopy>> lps;  // FP32 + refinement with host rescaling only
  int stream_main = 0, stream_fb = 1;        // ctx_stream numbers
  Profiler prof;                             // the FP32 (or plain FP64) engine
  Profiler fbprof;                           // the FP64 engine of stalled instances
  std::string err;
};

// A loaded batch: its LPs prepared on the host (validation, numbering,
// slices) and copied to the device, not yet rescaled or solved.  Instances
// are split by size into groups that run side by side: every launch of an
// engine carries one attempt of each of its running instances, so a small
// instance that needs many iterations would otherwise advance at the pace
// of the large instances' launches.
struct pdlp_batch {
  pdlp_ctx *ctx = nullptr;
  pdlp_params prm{};
  int32_t count = 0;
  std::unique_ptr<DevicePool> own_pool;  // no-allocation solves: the scratch region
                                         // (declared first: outlives the engines)
  std::vector<std::unique_ptr<BatchGroup>> groups;
  double load_sec = 0.0;
  bool ran = false;
  ~pdlp_batch() {
    for (auto &g : groups) {
      g->main.reset();
      for (hipEvent_t ev : g->prof.events) (void)hipEventDestroy(ev);
      for (hipEvent_t ev : g->fbprof.events) (void)hipEventDestroy(ev);
    }
  }
};

static int batch_group_count(int32_t count, const pdlp_params *params) {
  int g = params->batch_groups > 0 ? params->batch_groups : (count >= 24 ? 3 : 1);
  return std::max(1, std::min(g, std::min<int>(count, 8)));
}

// Device bytes an FP64 instance needs beyond its arena while it is prepared
// and solved (rescaling scratch, slice and block tables, partial sums, the
// step-size tables), an upper bound.
static size_t solve_overhead_bytes(const Prepared &P) {
  const size_t n = (size_t)P.n, m = (size_t)P.m, nnz = (size_t)P.nnz;
  const size_t slices = P.ps.sl.size() + P.ds.sl.size();
  const size_t blocks = 2 * (nnz / kBlockNnz + (n + m) / kBlockRows + 4);
  return 20 * m + 8 * n + 2 * slices * sizeof(SliceDesc) + 48 * slices +
         blocks * (sizeof(BlockDesc) + 2 * 8 * (kEvalN + 1)) + 2 * 8 * 65536 + (4 << 20);
}

static size_t footprint(const Prepared &P, int prec);

// pool: the device pool of the batch's engines (null: the context's);
// pageable: host staging without pinned memory; cap > 0: the LPs must fit
// in a pool of that many bytes (checked on the host before any device work).
static int load_batch(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps, const pdlp_params *params,
                      std::unique_ptr<pdlp_batch> *out, std::unique_ptr<DevicePool> pool = nullptr,
                      bool pageable = false, size_t cap = 0) {
  std::string *err = &ctx->err;
  const double t0 = now_sec();
  auto b = std::make_unique<pdlp_batch>();
  b->ctx = ctx;
  b->prm = *params;
  b->count = count;
  b->own_pool = std::move(pool);
  DevicePool *dpool = b->own_pool ? b->own_pool.get() : &ctx->pool;
  const bool refine = params->precision == PDLP_PRECISION_FP32_REFINE;
  const bool keep = refine && !(params->device_preconditioning && !params->l2_norm_rescaling);
  // groups of equal instance counts, in order of nonzeros
  const int G = batch_group_count(count, params);
  std::vector<int32_t> order(count);
  for (int32_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t c) { return lps[a].nnz < lps[c].nnz; });
  for (int g = 0; g < G; ++g) {
    auto grp = std::make_unique<BatchGroup>();
    grp->stream_main = 2 * g;
    grp->stream_fb = 2 * g + 1;
    hipStream_t s;
    int rc = ctx_stream(ctx, grp->stream_main, &s, err);
    if (rc) return rc;
    grp->main = std::make_unique<Engine>(s, dpool, *params, &grp->prof);
    grp->main->set_pageable_host(pageable);
    const int32_t lo = (int32_t)((int64_t)count * g / G), hi = (int32_t)((int64_t)count * (g + 1) / G);
    for (int32_t k = lo; k < hi; ++k) grp->idx.push_back(order[k]);
    std::sort(grp->idx.begin(), grp->idx.end());
    b->groups.push_back(std::move(grp));
  }
  // Host preparation (validation, numbering, slice tables) of the instances
  // on up to 16 host threads; then each is copied to the device in group
  // order on its group's stream.
  std::vector<std::unique_ptr<HostInst>> prepared(count);
  std::vector<std::string> perr(count);
  std::vector<int> prc(count, 0);
  std::atomic<int32_t> next{0};
  auto work = [&]() {
    for (int32_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
      auto h = std::make_unique<HostInst>();
      prc[i] = prepare_instance(&lps[i], params, h.get(), &perr[i]);
      prepared[i] = std::move(h);
    }
  };
  const int nth = std::max(1, std::min({(int)count, 16, (int)std::thread::hardware_concurrency()}));
  {
    std::vector<std::thread> threads;
    for (int t = 1; t < nth; ++t) threads.emplace_back(work);
    work();
    for (auto &t : threads) t.join();
  }
  if (cap > 0) {
    size_t need = 0;
    for (int32_t i = 0; i < count; ++i)
      if (!prc[i]) need += footprint(prepared[i]->P, 0) + solve_overhead_bytes(prepared[i]->P);
    if (need > cap) {
      *err = "the problem needs about " + std::to_string(need) + " bytes of device memory; " +
             "the no-allocation region holds " + std::to_string(cap);
      return 7;
    }
  }
  for (auto &grp : b->groups) {
    for (int32_t i : grp->idx) {
      int rc = prc[i];
      if (rc) *err = perr[i];
      else rc = grp->main->stage(std::move(prepared[i]), err);
      if (rc) {
        if (count > 1) *err = "instance " + std::to_string(i) + ": " + *err;
        return rc;
      }
      if (keep) grp->lps.push_back(std::make_unique<LpCopy>(lps[i]));
    }
  }
  b->load_sec = now_sec() - t0;
  *out = std::move(b);
  return 0;
}

// Solves one group of a loaded batch, filling results[idx[q]]; t0 is the
// start of the timed call (solve_time_sec counts from it).  FP32 +
// refinement: an instance whose refinement stalls is moved, as soon as it
// stalls, to a plain FP64 engine on the group's second stream, within what
// is left of the caller's iteration and time limits; the two engines'
// rounds are enqueued back to back so their kernels overlap on the GPU.  A
// moved instance's result is the FP64 solve's (or, if that solve stops at a
// caller's limit further from optimal, the refined point: RefinedPoint),
// with the FP32 iteration and attempt counts added and fp64_fallback = 1.
static int run_group(pdlp_batch *b, BatchGroup &grp, pdlp_result *results, pdlp_stats_cb cb,
                     void *user, double t0) {
  pdlp_ctx *ctx = b->ctx;
  std::string *err = &grp.err;
  Engine &main = *grp.main;
  const int32_t count = (int32_t)grp.idx.size();
  for (int32_t q = 0; q < count; ++q) {
    HostInst &h = main.inst((size_t)q);
    pdlp_result *r = &results[grp.idx[q]];
    h.res = r;
    h.t_submit = t0;
    h.cb = cb;
    h.user = user;
    r->refinement_rounds = 0;
    r->fp64_fallback = 0;
  }
  int rc = main.activate_staged(err);  // rescaling and slicing on the device
  if (rc) return rc;
  const double prep = b->load_sec + (now_sec() - t0);
  for (int32_t q = 0; q < count; ++q) results[grp.idx[q]].preprocess_time_sec = prep;
  const pdlp_params *params = &b->prm;
  if (params->precision != PDLP_PRECISION_FP32_REFINE) return main.run(err);

  pdlp_params p64 = *params;
  p64.precision = PDLP_PRECISION_FP64;
  hipStream_t s_fb;
  rc = ctx_stream(ctx, grp.stream_fb, &s_fb, err);  // created before the groups start
  if (rc) return rc;
  std::unique_ptr<Engine> fb;  // profiled into grp.fbprof
  struct Prior {
    int64_t iters, attempts;
    int32_t rounds;
  };
  std::vector<Prior> prior(count);
  std::vector<RefinedPoint> refined(count);
  std::vector<char> moved(count, 0);
  std::vector<EarlyFp64> early(count);
  const bool early_on = params->refine_early_fallback > 0 && !cb;
  bool main_done = false;
  for (;;) {
    const int r1 = main_done ? 1 : main.begin_round(err);
    if (r1 < 0) return r1;
    // the FP64 engine runs its (short) rounds while the FP32 chunk is in flight
    if (fb && !fb->all_done()) {
      do {
        rc = fb->step_round(err);
        if (rc < 0) return rc;
      } while (r1 == 0 && !main.round_ready() && !fb->all_done());
    }
    if (!main_done) {
      rc = main.end_round(err);
      if (rc < 0) return rc;
    }
    if (r1 == 1) main_done = true;
    for (int32_t q = 0; q < count; ++q) {
      const HostInst &h = main.inst((size_t)q);
      pdlp_result *r = &results[grp.idx[q]];
      EarlyFp64 &e = early[q];
      const double now = now_sec();
      // an early FP64 solve whose FP32 instance has left its round: keep it
      // if the refinement stalled and the solve is the one a start at the
      // stall would make, else cancel it (once it is active)
      if (e.k >= 0 && !e.committed && (h.done || h.ref_round != e.round)) {
        HostInst &g = fb->inst((size_t)e.k);
        const int64_t left_it = (int64_t)p64.termination_criteria.iteration_limit - r->iteration_count;
        const double left_t = p64.termination_criteria.time_sec_limit - (now - t0);
        // with an iteration limit of left_it a solve evaluates at the same
        // iterations up to left_it + 1, so it is the same solve as long as
        // its next (or final) evaluation is not past that point
        if (h.done && h.stalled && !moved[q] && left_it > 0 && left_t > 0 &&
            g.eval_iter <= left_it + 1) {
          g.iteration_limit_override = left_it;
          g.time_limit_override = (g.staged ? 0.0 : now - g.start_time) + left_t;
          e.committed = true;
          e.commit_sec = now - t0;
          moved[q] = 1;
          prior[q] = {r->iteration_count, r->total_attempts, r->refinement_rounds};
          refined[q].save(*r, h.P.n, h.P.m);
          r->fp64_fallback = 2;
        } else if (!g.staged) {
          rc = fb->cancel((size_t)e.k, err);
          if (rc) return rc;
          e.k = -1;
        }
      }
      if (early_on && e.k < 0 && !h.done && h.ref_round >= 1 && !r->x_current && !r->y_current) {
        const int64_t at = params->refine_early_fallback;
        const int64_t sofar = h.iters_total + h.iter_start - 1;
        const int64_t left_it = (int64_t)p64.termination_criteria.iteration_limit - sofar;
        const double left_t = p64.termination_criteria.time_sec_limit - (now - t0);
        if (h.inner_cap > at && h.iter_start - 1 >= at && left_it > 0 && left_t > 0) {
          if (!fb) fb = std::make_unique<Engine>(s_fb, &ctx->pool, p64, &grp.fbprof);
          e.x.assign((size_t)std::max<int64_t>(h.P.n, 1), 0.0);
          e.y.assign((size_t)std::max<int64_t>(h.P.m, 1), 0.0);
          e.res = pdlp_result{};
          e.res.x = e.x.data();
          e.res.y = e.y.data();
          auto g = std::make_unique<HostInst>();
          g->res = &e.res;
          g->t_submit = t0;
          g->iteration_limit_override = left_it;
          g->time_limit_override = left_t;
          const int32_t k = (int32_t)fb->count();
          rc = stage_fp64(fb.get(), h, grp.lps.empty() ? nullptr : grp.lps[q].get(), p64,
                          std::move(g), err);
          if (rc) return rc;
          e.k = k;
          e.round = h.ref_round;
        }
      }
      if (!h.done || !h.stalled || moved[q]) continue;
      moved[q] = 1;
      const int64_t left_it = (int64_t)p64.termination_criteria.iteration_limit - r->iteration_count;
      const double left_t = p64.termination_criteria.time_sec_limit - (now - t0);
      if (left_it <= 0 || !(left_t > 0)) continue;
      prior[q] = {r->iteration_count, r->total_attempts, r->refinement_rounds};
      refined[q].save(*r, h.P.n, h.P.m);
      if (!fb) fb = std::make_unique<Engine>(s_fb, &ctx->pool, p64, &grp.fbprof);
      auto g = std::make_unique<HostInst>();
      g->res = r;
      g->t_submit = t0;
      g->cb = cb;
      g->user = user;
      g->iteration_limit_override = left_it;
      g->time_limit_override = left_t;
      rc = stage_fp64(fb.get(), h, grp.lps.empty() ? nullptr : grp.lps[q].get(), p64,
                      std::move(g), err);
      if (rc) return rc;
      r->fp64_fallback = 2;  // being solved in FP64
    }
    bool pending = false;  // early solves still to be cancelled
    for (const EarlyFp64 &e : early) pending = pending || (e.k >= 0 && !e.committed);
    if (main_done && !pending && (!fb || fb->all_done())) break;
  }
  for (int32_t q = 0; q < count; ++q) {
    pdlp_result *r = &results[grp.idx[q]];
    const EarlyFp64 &e = early[q];
    if (e.committed) {
      // the fields a plain FP64 solve writes into its result
      const Prepared &P = main.inst((size_t)q).P;
      if (P.n) memcpy(r->x, e.x.data(), 8 * (size_t)P.n);
      if (P.m) memcpy(r->y, e.y.data(), 8 * (size_t)P.m);
      r->termination_reason = e.res.termination_reason;
      // the result exists once the refinement has stalled
      r->solve_time_sec = std::max(e.res.solve_time_sec, e.commit_sec);
      r->iteration_count = e.res.iteration_count;
      r->final_stats = e.res.final_stats;
      r->total_attempts = e.res.total_attempts;
      r->refinement_rounds = e.res.refinement_rounds;
    }
    if (r->fp64_fallback != 2) continue;
    const int64_t it = (int64_t)r->iteration_count + prior[q].iters;
    r->iteration_count = (int32_t)std::min<int64_t>(it, INT32_MAX);
    r->total_attempts += prior[q].attempts;
    r->refinement_rounds = prior[q].rounds;
    r->fp64_fallback = 1;
    refined[q].restore_if_better(r, params->termination_criteria.optimality_norm);
  }
  return 0;
}

// Solves a loaded batch: group 0 on the calling thread, the others on host
// threads of their own (each drives its engines on its own streams).
static int run_batch(pdlp_batch *b, pdlp_result *results, pdlp_stats_cb cb, void *user,
                     double t0) {
  pdlp_ctx *ctx = b->ctx;
  if (b->ran) {
    ctx->err = "a loaded batch is solved once";
    return 1;
  }
  b->ran = true;
  const int G = (int)b->groups.size();
  // kernel timing creates HIP events: not in a no-allocation solve
  const bool prof_on = ctx->prof.enabled && !b->own_pool;
  for (auto &grp : b->groups) {
    grp->prof.enabled = grp->fbprof.enabled = prof_on;
    grp->fbprof.timed = ctx->prof_fallback_timed;
  }
  if (b->prm.precision == PDLP_PRECISION_FP32_REFINE) {
    for (auto &grp : b->groups) {
      hipStream_t s;
      int rc = ctx_stream(ctx, grp->stream_fb, &s, &ctx->err);
      if (rc) return rc;
    }
  }
  std::vector<int> rcs(G, 0);
  std::vector<std::thread> threads;
  for (int g = 1; g < G; ++g)
    threads.emplace_back([&, g]() {
      (void)hipSetDevice(ctx->device);
      rcs[g] = run_group(b, *b->groups[g], results, cb, user, t0);
    });
  if (G > 0) rcs[0] = run_group(b, *b->groups[0], results, cb, user, t0);
  for (auto &t : threads) t.join();
  pdlp_kernel_profile &acc = ctx->prof.acc;
  for (auto &grp : b->groups) {
    const pdlp_kernel_profile &p = grp->prof.acc;
    acc.primal_launches += p.primal_launches;
    acc.dual_launches += p.dual_launches;
    acc.primal_ms += p.primal_ms;
    acc.dual_ms += p.dual_ms;
    acc.primal_bytes += p.primal_bytes;
    acc.dual_bytes += p.dual_bytes;
    const pdlp_kernel_profile &f = grp->fbprof.acc;
    acc.fallback_primal_launches += f.primal_launches;
    acc.fallback_dual_launches += f.dual_launches;
    acc.fallback_primal_ms += f.primal_ms;
    acc.fallback_dual_ms += f.dual_ms;
    acc.fallback_primal_bytes += f.primal_bytes;
    acc.fallback_dual_bytes += f.dual_bytes;
    grp->prof.acc = pdlp_kernel_profile{};
    grp->fbprof.acc = pdlp_kernel_profile{};
  }
  for (int g = 0; g < G; ++g)
    if (rcs[g]) {
      ctx->err = b->groups[g]->err;
      return rcs[g];
    }
  return 0;
}

extern "C" int pdlp_solve(pdlp_ctx *ctx, const pdlp_lp *lp, const pdlp_params *params,
                          pdlp_result *result, pdlp_stats_cb cb, void *user) {
  if (!ctx) return 1;
  ctx->err.clear();
  if (!lp || !params || !result) {
    ctx->err = "null argument";
    return 1;
  }
  if ((lp->n > 0 && !result->x) || (lp->m > 0 && !result->y)) {
    ctx->err = "result vectors must be allocated by the caller";
    return 1;
  }
  const double t0 = now_sec();
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  std::unique_ptr<pdlp_batch> b;
  rc = load_batch(ctx, 1, lp, params, &b);
  if (rc == 0) rc = run_batch(b.get(), result, cb, user, t0);
  if (rc) return rc;
  result->solve_time_sec = now_sec() - t0;
  return 0;
}

extern "C" int pdlp_solve_noalloc(pdlp_ctx *ctx, const pdlp_lp *lp, const pdlp_params *params,
                                  pdlp_result *result, pdlp_stats_cb cb, void *user) {
  if (!ctx) return 1;
  ctx->err.clear();
  if (!lp || !params || !result) {
    ctx->err = "null argument";
    return 1;
  }
  if ((lp->n > 0 && !result->x) || (lp->m > 0 && !result->y)) {
    ctx->err = "result vectors must be allocated by the caller";
    return 1;
  }
  const double t0 = now_sec();
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  if (params->precision != PDLP_PRECISION_FP64) {
    ctx->err = "pdlp_solve_noalloc solves in PDLP_PRECISION_FP64 only";
    return 2;
  }
  if (ctx->scratch_rc) {
    ctx->err = ctx->scratch_err;
    return ctx->scratch_rc;
  }
  (void)hipSetDevice(ctx->device);
  void *base = scratch_region(ctx->device);
  if (!base) {
    ctx->err = "no-allocation region not set up on this device";
    return -1;
  }
  std::unique_ptr<pdlp_batch> b;
  rc = load_batch(ctx, 1, lp, params, &b, std::make_unique<DevicePool>(base, kScratchBytes),
                  true, kScratchBytes);
  if (rc == 0) rc = run_batch(b.get(), result, cb, user, t0);
  if (rc < 0 && b && b->own_pool && ctx->err.find("allocation") != std::string::npos) {
    // the region filled up beyond the host-side estimate
    ctx->err = "the problem does not fit the no-allocation region of " +
               std::to_string(kScratchBytes) + " bytes (" + ctx->err + ")";
    rc = 7;
  }
  if (rc) return rc;
  result->solve_time_sec = now_sec() - t0;
  return 0;
}

extern "C" int pdlp_batch_solve(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps,
                                const pdlp_params *params, pdlp_result *results) {
  if (!ctx) return 1;
  ctx->err.clear();
  if (count < 0 || (count > 0 && (!lps || !results)) || !params) {
    ctx->err = "invalid batch arguments";
    return 1;
  }
  const double t0 = now_sec();
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  for (int32_t i = 0; i < count; ++i) {
    if ((lps[i].n > 0 && !results[i].x) || (lps[i].m > 0 && !results[i].y)) {
      ctx->err = "result vectors must be allocated by the caller (instance " + std::to_string(i) + ")";
      return 1;
    }
  }
  std::unique_ptr<pdlp_batch> b;
  rc = load_batch(ctx, count, lps, params, &b);
  if (rc == 0) rc = run_batch(b.get(), results, nullptr, nullptr, t0);
  return rc;
}

// A candidate replaces the current choice only when at least 3 % faster:
// repeated trials of one configuration differ by 1-2 % on the 512-LP batch,
// and a smaller margin switched between near-equal choices from run to run.
constexpr double kTuneGain = 0.97;

extern "C" int pdlp_tune(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps,
                         const pdlp_params *params, int32_t trial_iterations, pdlp_tuning *out) {
  if (!ctx) return 1;
  ctx->err.clear();
  if (!out || !params || count <= 0 || !lps || trial_iterations <= 0) {
    ctx->err = "invalid tuning arguments";
    return 1;
  }
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  int64_t sn = 0, sm = 0, snz = 0;
  for (int32_t i = 0; i < count; ++i) {
    sn += lps[i].n;
    sm += lps[i].m;
    snz += lps[i].nnz;
  }
  const std::string key = std::to_string(count) + "/" + std::to_string(sn) + "/" +
                          std::to_string(sm) + "/" + std::to_string(snz) + "/" +
                          std::to_string(params->precision) + "/" +
                          std::to_string(batch_group_count(count, params));
  auto hit = ctx->tune_cache.find(key);
  if (hit != ctx->tune_cache.end()) {
    *out = hit->second;
    out->candidates = 0;
    out->seconds = 0.0;
    return 0;
  }
  (void)hipSetDevice(ctx->device);
  const double t0 = now_sec();
  std::vector<std::vector<double>> xs(count), ys(count);
  std::vector<pdlp_result> res(count);
  for (int32_t i = 0; i < count; ++i) {
    xs[i].assign((size_t)std::max<int64_t>(lps[i].n, 1), 0.0);
    ys[i].assign((size_t)std::max<int64_t>(lps[i].m, 1), 0.0);
    memset(&res[i], 0, sizeof res[i]);
    res[i].x = xs[i].data();
    res[i].y = ys[i].data();
  }
  int ntrials = 0;
  // k_primal and k_dual time per launch over trial_iterations iterations
  // with tolerances no point meets, timed with the kernels' own events
  auto trial = [&](int lb, int u, int w, int po, int dov, double *pus, double *dus) -> int {
    pdlp_params p = *params;
    p.lane_budget = lb;
    p.slice_unroll = u;
    p.slice_waves = w;
    p.primal_occupancy = po;
    p.dual_occupancy = dov;
    p.record_iteration_stats = 0;
    p.termination_criteria.iteration_limit = trial_iterations;
    p.termination_criteria.eps_optimal_absolute = 1e-300;
    p.termination_criteria.eps_optimal_relative = 1e-300;
    p.termination_criteria.time_sec_limit = INFINITY;
    p.termination_criteria.kkt_matrix_pass_limit = INFINITY;
    std::unique_ptr<pdlp_batch> b;
    int r = load_batch(ctx, count, lps, &p, &b);
    if (r) return r;
    const pdlp_kernel_profile saved = ctx->prof.acc;
    const bool saved_on = ctx->prof.enabled, saved_fb = ctx->prof_fallback_timed;
    ctx->prof.enabled = true;
    ctx->prof_fallback_timed = false;
    memset(&ctx->prof.acc, 0, sizeof ctx->prof.acc);
    r = run_batch(b.get(), res.data(), nullptr, nullptr, now_sec());
    const pdlp_kernel_profile got = ctx->prof.acc;
    ctx->prof.acc = saved;
    ctx->prof.enabled = saved_on;
    ctx->prof_fallback_timed = saved_fb;
    if (r) return r;
    ++ntrials;
    const double nl = (double)got.primal_launches;
    *pus = nl > 0 ? got.primal_ms / nl * 1e3 : INFINITY;
    *dus = nl > 0 ? got.dual_ms / nl * 1e3 : INFINITY;
    return 0;
  };
  pdlp_tuning best{};
  best.lane_budget = 64;
  best.slice_unroll = 0;  // each kernel's default
  best.slice_waves = 4;
  double bp = 0.0, bd = 0.0;  // per-kernel times of the current best
  rc = trial(64, 0, 4, 0, 0, &bp, &bd);
  if (rc) return rc;
  best.kernel_us_per_iteration = bp + bd;
  best.default_us_per_iteration = best.kernel_us_per_iteration;
  auto consider = [&](int lb, int u, int w) -> int {
    double pus = INFINITY, dus = INFINITY;
    int r = trial(lb, u, w, 0, 0, &pus, &dus);
    if (r) return r;
    if (pus + dus < kTuneGain * best.kernel_us_per_iteration) {
      best.lane_budget = lb;
      best.slice_unroll = u;
      best.slice_waves = w;
      best.kernel_us_per_iteration = pus + dus;
      bp = pus;
      bd = dus;
    }
    return 0;
  };
  for (int lb : {32, 128})
    if ((rc = consider(lb, 0, 4))) return rc;
  if ((rc = consider(best.lane_budget, 0, 2))) return rc;
  if ((rc = consider(best.lane_budget, 4, best.slice_waves))) return rc;
  // FP64: 8 for the dual kernel too (the default gives it 16)
  if (params->precision != PDLP_PRECISION_FP32_REFINE && best.slice_unroll == 0 &&
      (rc = consider(best.lane_budget, 8, best.slice_waves)))
    return rc;
  // occupancy caps, each kernel by its own launches (0 = the default caps)
  int po = 0, dov = 0;
  for (int occ : {3, 4, 5, 6, 8}) {
    double pus = INFINITY, dus = INFINITY;
    if ((rc = trial(best.lane_budget, best.slice_unroll, best.slice_waves, occ, occ, &pus, &dus)))
      return rc;
    if (pus < kTuneGain * bp) bp = pus, po = occ;
    if (dus < kTuneGain * bd) bd = dus, dov = occ;
  }
  best.primal_occupancy = po;
  best.dual_occupancy = dov;
  best.kernel_us_per_iteration = bp + bd;
  best.candidates = ntrials;
  best.seconds = now_sec() - t0;
  ctx->tune_cache[key] = best;
  *out = best;
  return 0;
}

extern "C" int pdlp_batch_load(pdlp_ctx *ctx, int32_t count, const pdlp_lp *lps,
                               const pdlp_params *params, pdlp_batch **out) {
  if (!ctx) return 1;
  ctx->err.clear();
  if (!out || count < 0 || (count > 0 && !lps) || !params) {
    ctx->err = "invalid batch arguments";
    return 1;
  }
  *out = nullptr;
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  std::unique_ptr<pdlp_batch> b;
  rc = load_batch(ctx, count, lps, params, &b);
  if (rc) return rc;
  *out = b.release();
  return 0;
}

extern "C" int pdlp_batch_run(pdlp_batch *batch, pdlp_result *results) {
  if (!batch) return 1;
  pdlp_ctx *ctx = batch->ctx;
  ctx->err.clear();
  if (batch->count > 0 && !results) {
    ctx->err = "invalid batch arguments";
    return 1;
  }
  for (auto &grp : batch->groups)
    for (size_t q = 0; q < grp->idx.size(); ++q) {
      const Prepared &P = grp->main->inst(q).P;
      const int32_t i = grp->idx[q];
      if ((P.n > 0 && !results[i].x) || (P.m > 0 && !results[i].y)) {
        ctx->err = "result vectors must be allocated by the caller (instance " + std::to_string(i) + ")";
        return 1;
      }
    }
  (void)hipSetDevice(ctx->device);
  return run_batch(batch, results, nullptr, nullptr, now_sec());
}

extern "C" void pdlp_batch_free(pdlp_batch *batch) {
  if (!batch) return;
  (void)hipSetDevice(batch->ctx->device);
  delete batch;
}

// ------------------------------------------------------------------ serving
struct ServeReq {
  int64_t id;
  std::unique_ptr<HostInst> inst;
  double tol;
  size_t bytes;
  std::unique_ptr<LpCopy> lp;
  double time_limit = INFINITY;  // the request's own, from submission
  int64_t iter_limit = -1;       // the request's own (< 0: none)
};

struct ServeDone {
  std::vector<double> x, y;
  int64_t n = 0, m = 0;
  pdlp_result res;
  std::unique_ptr<LpCopy> lp;  // FP32 + refinement servers with host rescaling
  double tol = 0.0, t_submit = 0.0;
  double time_limit = INFINITY;  // the request's own limits (see ServeReq)
  int64_t iter_limit = -1;
  int64_t prior_iters = 0, prior_attempts = 0;
  int32_t prior_rounds = 0;
  size_t fb_bytes = 0;         // footprint of the FP64 solve, while it runs
  RefinedPoint refined;        // the stalled FP32 solve's point
};

// Serving size classes: a request goes to the engine of its class by
// nonzeros (class c holds nnz in [2^(17+3(c-1)), 2^(17+3c)), the last one
// everything above), so a small request does not advance at the pace of the
// large requests' launches; the engines run their rounds side by side on
// streams of their own.  params.batch_groups sets the number of classes
// (0 = 3, at most kServeClasses); results do not depend on it.
constexpr int kServeClasses = 4;

static int serve_class(int64_t nnz, int classes) {
  int c = 0;
  while (c + 1 < classes && nnz >= (int64_t{1} << (17 + 3 * c))) ++c;
  return c;
}

struct pdlp_server {
  pdlp_ctx *ctx = nullptr;
  pdlp_params prm{};
  double budget = 0;
  int32_t max_slots = 0;
  int classes = 1;
  std::vector<std::unique_ptr<Engine>> engs;  // one per size class, created on first use
  std::unique_ptr<Engine> fb;  // FP64 solves of stalled FP32 requests (context's second stream)
  std::deque<ServeReq> queue;
  std::map<int64_t, std::unique_ptr<ServeDone>> running;  // result storage of admitted ids
  std::map<int64_t, std::unique_ptr<ServeDone>> done;
  std::vector<int64_t> finished;
  int64_t next_id = 0;
  double in_use = 0;
};

static size_t footprint(const Prepared &P, int prec) {
  const size_t ts = prec ? 4 : 8;
  size_t b = (size_t)P.nnz * 2 * (4 + ts) + (size_t)(P.n + P.m + 2) * 4 + ts * (14 * P.n + 11 * P.m) +
             8 * (4 * P.n + 2 * P.m) + 64 * 1024;
  b += (size_t)(P.ps.slots + P.ds.slots) * (4 + ts);  // sliced copies of A^T and A
  // FP32 + refinement: the FP64 matrix twice, ylo (T and FP64) and the
  // refinement vectors (6 n-vectors, 4 m-vectors)
  if (prec) b += (size_t)P.nnz * 16 + (ts + 8) * P.m + 8 * (6 * P.n + 4 * P.m) + 4 * (2 * P.n + 4 * P.m);
  return b;
}

extern "C" int pdlp_serve_open(pdlp_ctx *ctx, const pdlp_params *params, double hbm_budget_bytes,
                               int32_t max_slots, pdlp_server **out) {
  if (!ctx || !params || !out) return 1;
  ctx->err.clear();
  int rc = check_params(params, &ctx->err);
  if (rc) return rc;
  auto *s = new pdlp_server();
  s->ctx = ctx;
  s->prm = *params;
  s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : 1e18;
  s->max_slots = max_slots > 0 ? max_slots : 1 << 30;
  s->classes = params->batch_groups > 0 ? std::min(params->batch_groups, kServeClasses) : 3;
  s->engs.resize(s->classes);
  *out = s;
  return 0;
}

// The engine of size class c (its stream: the context's first for class 0,
// a stream of its own for the others).
static int serve_engine(pdlp_server *s, int c, Engine **out, std::string *err) {
  if (!s->engs[c]) {
    hipStream_t st;
    int rc = ctx_stream(s->ctx, c == 0 ? 0 : 16 + c, &st, err);
    if (rc) return rc;
    s->engs[c] = std::make_unique<Engine>(st, &s->ctx->pool, s->prm);
  }
  *out = s->engs[c].get();
  return 0;
}

static int32_t serve_running(const pdlp_server *s) {
  size_t k = 0;
  for (const auto &e : s->engs)
    if (e) k += e->count();
  return (int32_t)std::min<size_t>(k, INT32_MAX);
}

static int serve_submit(pdlp_server *s, const pdlp_lp *lp, double tolerance, double time_limit,
                        int64_t iter_limit, int64_t *id) {
  if (!s || !lp || !id) return 1;
  auto h = std::make_unique<HostInst>();
  int rc = pdlp::prepare_instance(lp, &s->prm, h.get(), &s->ctx->err);
  if (rc) return rc;
  h->tolerance_override = tolerance;
  h->t_submit = now_sec();
  ServeReq r;
  r.id = s->next_id++;
  r.tol = tolerance;
  r.time_limit = time_limit > 0.0 ? time_limit : INFINITY;
  r.iter_limit = iter_limit;
  r.bytes = footprint(h->P, s->prm.precision == PDLP_PRECISION_FP32_REFINE);
  if ((double)r.bytes > s->budget) {
    s->ctx->err = "request exceeds the HBM budget";
    return 5;
  }
  h->tag = r.id;
  r.inst = std::move(h);
  // an FP32 request whose refinement stalls is solved again in FP64: from
  // its own unscaled device copies with device rescaling, else from a host
  // copy of its LP
  if (s->prm.precision == PDLP_PRECISION_FP32_REFINE &&
      !(s->prm.device_preconditioning && !s->prm.l2_norm_rescaling))
    r.lp = std::make_unique<LpCopy>(*lp);
  *id = r.id;
  s->queue.push_back(std::move(r));
  return 0;
}

extern "C" int pdlp_serve_submit(pdlp_server *s, const pdlp_lp *lp, double tolerance, int64_t *id) {
  return serve_submit(s, lp, tolerance, INFINITY, -1, id);
}

extern "C" int pdlp_serve_submit_limits(pdlp_server *s, const pdlp_lp *lp, double tolerance,
                                        double time_sec_limit, int32_t iteration_limit,
                                        int64_t *id) {
  return serve_submit(s, lp, tolerance, time_sec_limit, iteration_limit, id);
}

// Result storage of a request leaving the queue (admitted, expired or cancelled).
static std::unique_ptr<ServeDone> serve_record(ServeReq &r) {
  auto d = std::make_unique<ServeDone>();
  d->x.assign(std::max<int64_t>(r.inst->P.n, 1), 0.0);
  d->y.assign(std::max<int64_t>(r.inst->P.m, 1), 0.0);
  d->n = r.inst->P.n;
  d->m = r.inst->P.m;
  memset(&d->res, 0, sizeof d->res);
  d->res.x = d->x.data();
  d->res.y = d->y.data();
  d->lp = std::move(r.lp);
  d->tol = r.tol;
  d->t_submit = r.inst->t_submit;
  d->time_limit = r.time_limit;
  d->iter_limit = r.iter_limit;
  return d;
}

// A request finished without a solution (its deadline passed in the queue,
// or it was cancelled): reason, zero x and y.
static void serve_finish_empty(pdlp_server *s, int64_t id, std::unique_ptr<ServeDone> d,
                               int32_t reason) {
  d->res.termination_reason = reason;
  d->res.solve_time_sec = now_sec() - d->t_submit;
  std::fill(d->x.begin(), d->x.end(), 0.0);
  std::fill(d->y.begin(), d->y.end(), 0.0);
  s->done[id] = std::move(d);
  s->finished.push_back(id);
}

extern "C" int pdlp_serve_poll(pdlp_server *s, int64_t *ids, int32_t max_ids, int32_t *finished) {
  if (!s || !finished) return 1;
  std::string *err = &s->ctx->err;
  (void)hipSetDevice(s->ctx->device);
  // requests whose own time limit ran out while queued finish with TIME_LIMIT
  const double t_now = now_sec();
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    if (t_now - it->inst->t_submit >= it->time_limit) {
      const int64_t id = it->id;
      serve_finish_empty(s, id, serve_record(*it), PDLP_TERMINATION_REASON_TIME_LIMIT);
      it = s->queue.erase(it);
    } else {
      ++it;
    }
  }
  // admit by footprint and slot count
  while (!s->queue.empty() && serve_running(s) < s->max_slots &&
         s->in_use + (double)s->queue.front().bytes <= s->budget) {
    ServeReq r = std::move(s->queue.front());
    s->queue.pop_front();
    auto d = serve_record(r);
    r.inst->res = &d->res;
    // the request's own limits on top of the server's; its time counts from
    // submission (the solve's clock starts at admission)
    if (r.iter_limit >= 0) r.inst->iteration_limit_override = r.iter_limit;
    if (std::isfinite(r.time_limit))
      r.inst->time_limit_override =
          std::max(0.0, r.time_limit - (now_sec() - r.inst->t_submit));
    // counted as running only once the engine holds it; a request the
    // engine cannot take is finished with reason OTHER (its result says so)
    // and the error is returned
    Engine *eng = nullptr;
    int rc = serve_engine(s, serve_class(r.inst->P.nnz, s->classes), &eng, err);
    if (rc == 0) rc = eng->add(std::move(r.inst), err);
    if (rc) {
      d->res.termination_reason = PDLP_TERMINATION_REASON_OTHER;
      d->res.solve_time_sec = now_sec() - d->t_submit;
      s->done[r.id] = std::move(d);
      s->finished.push_back(r.id);
      return rc;
    }
    s->in_use += (double)r.bytes;
    s->running[r.id] = std::move(d);
  }
  const bool refine = s->prm.precision == PDLP_PRECISION_FP32_REFINE;
  // one round of every engine, enqueued back to back so their kernels overlap
  std::vector<Engine *> active;
  for (auto &e : s->engs)
    if (e && e->count() > 0) active.push_back(e.get());
  if (s->fb && s->fb->count() > 0) active.push_back(s->fb.get());
  for (Engine *e : active) {
    int rc = e->begin_round(err);
    if (rc < 0) return rc;
  }
  for (Engine *e : active) {
    int rc = e->end_round(err);
    if (rc < 0) return rc;
  }
  for (auto &ep : s->engs) {
    if (!ep || ep->count() == 0) continue;
    Engine *eng = ep.get();
    int rc = 0;
    for (size_t q = 0; q < eng->count(); ++q) {
      HostInst &h = eng->inst(q);
      if (!h.done) continue;
      auto it = s->running.find(h.tag);
      if (it == s->running.end()) continue;
      ServeDone &d = *it->second;
      s->in_use -= (double)footprint(h.P, refine);
      if (refine && h.stalled) {
        // refinement stalled: the request continues as a plain FP64 solve
        // within what is left of its iteration and time limits
        pdlp_params p64 = s->prm;
        p64.precision = PDLP_PRECISION_FP64;
        int64_t lim_it = (int64_t)p64.termination_criteria.iteration_limit;
        if (d.iter_limit >= 0) lim_it = std::min(lim_it, d.iter_limit);
        const double lim_t = std::min(p64.termination_criteria.time_sec_limit, d.time_limit);
        const int64_t left_it = lim_it - d.res.iteration_count;
        const double left_t = lim_t - (now_sec() - d.t_submit);
        if (left_it > 0 && left_t > 0) {
          auto g = std::make_unique<HostInst>();
          g->tolerance_override = d.tol;
          g->t_submit = d.t_submit;
          g->res = &d.res;
          g->tag = h.tag;
          g->iteration_limit_override = left_it;
          g->time_limit_override = left_t;
          d.prior_iters = d.res.iteration_count;
          d.prior_attempts = d.res.total_attempts;
          d.prior_rounds = d.res.refinement_rounds;
          d.refined.save(d.res, d.n, d.m);
          d.fb_bytes = footprint(h.P, 0);
          if (!s->fb) {
            hipStream_t s2;
            rc = ctx_stream(s->ctx, 1, &s2, err);
            if (rc) return rc;
            s->fb = std::make_unique<Engine>(s2, &s->ctx->pool, p64);
          }
          s->in_use += (double)d.fb_bytes;
          // staging copies on the device and waits for them, so the FP32
          // instance's arena can be recycled below
          rc = stage_fp64(s->fb.get(), h, d.lp.get(), p64, std::move(g), err);
          if (rc) return rc;
          continue;
        }
      }
      s->done[h.tag] = std::move(it->second);
      s->running.erase(it);
      s->finished.push_back(h.tag);
    }
    eng->remove_done();
  }
  if (s->fb && s->fb->count() > 0) {
    for (size_t q = 0; q < s->fb->count(); ++q) {
      HostInst &h = s->fb->inst(q);
      if (!h.done) continue;
      auto it = s->running.find(h.tag);
      if (it == s->running.end()) continue;
      ServeDone &d = *it->second;
      pdlp_result &r = d.res;
      const int64_t total = (int64_t)r.iteration_count + d.prior_iters;
      r.iteration_count = (int32_t)std::min<int64_t>(total, INT32_MAX);
      r.total_attempts += d.prior_attempts;
      r.refinement_rounds = d.prior_rounds;
      r.fp64_fallback = 1;
      d.refined.restore_if_better(&r, s->prm.termination_criteria.optimality_norm);
      d.refined = RefinedPoint();
      s->in_use -= (double)d.fb_bytes;
      s->done[h.tag] = std::move(it->second);
      s->running.erase(it);
      s->finished.push_back(h.tag);
    }
    s->fb->remove_done();
  }
  int32_t k = 0;
  while (k < max_ids && !s->finished.empty()) {
    ids[k++] = s->finished.front();
    s->finished.erase(s->finished.begin());
  }
  *finished = k;
  return 0;
}

extern "C" int pdlp_serve_result(pdlp_server *s, int64_t id, pdlp_result *result) {
  if (!s || !result) return 1;
  auto it = s->done.find(id);
  if (it == s->done.end()) {
    s->ctx->err = "request not finished or unknown";
    return 6;
  }
  ServeDone &d = *it->second;
  double *x = result->x, *y = result->y;
  double *xc = result->x_current, *yc = result->y_current;
  *result = d.res;
  result->x = x;
  result->y = y;
  result->x_current = xc;
  result->y_current = yc;
  if (x && d.n > 0) memcpy(x, d.x.data(), sizeof(double) * (size_t)d.n);
  if (y && d.m > 0) memcpy(y, d.y.data(), sizeof(double) * (size_t)d.m);
  s->done.erase(it);
  return 0;
}

extern "C" int pdlp_serve_cancel(pdlp_server *s, int64_t id) {
  if (!s) return 1;
  std::string *err = &s->ctx->err;
  (void)hipSetDevice(s->ctx->device);
  for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
    if (it->id != id) continue;
    serve_finish_empty(s, id, serve_record(*it), PDLP_TERMINATION_REASON_OTHER);
    s->queue.erase(it);
    return 0;
  }
  auto rit = s->running.find(id);
  if (rit == s->running.end()) {
    *err = "request not running or queued";
    return 6;
  }
  ServeDone &d = *rit->second;
  const bool refine = s->prm.precision == PDLP_PRECISION_FP32_REFINE;
  // the request is on its class's engine, or (a stalled FP32 request) on the FP64 one
  std::vector<Engine *> engines;
  for (auto &e : s->engs) engines.push_back(e.get());
  engines.push_back(s->fb.get());
  for (size_t ei = 0; ei < engines.size(); ++ei) {
    Engine *e = engines[ei];
    const int which = ei + 1 == engines.size() ? 1 : 0;
    if (!e) continue;
    for (size_t q = 0; q < e->count(); ++q) {
      HostInst &h = e->inst(q);
      if (h.tag != id || h.done) continue;
      const int64_t it_now = which == 0 && refine ? h.iters_total + h.iter_start - 1
                                                  : h.iter_start - 1;
      const int64_t att = (which == 0 && refine ? h.attempts_total : 0) + e->attempts(q);
      int rc = e->cancel(q, err);
      if (rc) return rc;
      s->in_use -= which == 0 ? (double)footprint(h.P, refine) : (double)d.fb_bytes;
      d.res.iteration_count = (int32_t)std::min<int64_t>(
          (which == 1 ? d.prior_iters : 0) + it_now, INT32_MAX);
      d.res.total_attempts = (which == 1 ? d.prior_attempts : 0) + att;
      d.res.refinement_rounds = which == 1 ? d.prior_rounds : h.ref_round;
      d.res.fp64_fallback = which == 1 ? 1 : 0;
      std::unique_ptr<ServeDone> rec = std::move(rit->second);
      s->running.erase(rit);
      serve_finish_empty(s, id, std::move(rec), PDLP_TERMINATION_REASON_OTHER);
      return 0;
    }
  }
  // admitted and finished in the engine, awaiting its FP64 staging: not expected
  *err = "request not running or queued";
  return 6;
}

extern "C" int pdlp_serve_close(pdlp_server *s) {
  if (!s) return 1;
  s->fb.reset();
  s->engs.clear();
  delete s;
  return 0;
}
