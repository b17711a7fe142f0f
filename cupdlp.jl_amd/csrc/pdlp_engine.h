// pdlp_engine.h — the host driver of the restarted-PDHG engine.
//
// An Engine owns a packed set of LP instances on one HIP stream.  It runs the
// reference's optimize() loop (src/primal_dual_hybrid_gradient_gpu.jl:457-798)
// for every instance at once: step attempts run on the device without host
// synchronisation; the host takes over only at the evaluation points
// (iterations 1..10, then every termination_evaluation_frequency), where it
// reads back a few dozen reduced scalars per instance and applies the
// termination test (src/termination.jl:206-246), the restart scheme
// (src/saddle_point_gpu.jl:479-622) and the primal-weight update (:627-654).
#pragma once
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "pdlp_device.h"
#include "pdlp_hip.h"
#include "pdlp_prep.h"

namespace pdlp {

// Caching device allocator owned by a context; the engines of a batch's
// groups share it from their host threads.
class DevicePool {
 public:
  DevicePool() = default;
  // A pool confined to a fixed device region [base, base + bytes): get()
  // carves 256-byte aligned blocks out of it (first fit) and returns null
  // when it is full; put() gives them back (adjacent free blocks merge).
  // It never calls hipMalloc or hipFree.
  DevicePool(void *base, size_t bytes);
  ~DevicePool();
  void *get(size_t bytes);
  void put(void *p);
  void release_all();
  size_t in_use() const {
    std::lock_guard<std::mutex> lk(mu_);
    return in_use_;
  }
  size_t peak() const {
    std::lock_guard<std::mutex> lk(mu_);
    return peak_;
  }

 private:
  struct Blk { void *p; size_t bytes; };
  mutable std::mutex mu_;
  std::vector<Blk> free_, used_;
  size_t in_use_ = 0, peak_ = 0;
  char *region_ = nullptr;  // fixed-region mode
  size_t region_bytes_ = 0;
};

struct HipError {
  hipError_t code;
  std::string where;
};

// The original problem data of an instance already resident on this device
// (an FP32 + refinement instance keeps its unscaled FP64 matrix, t_orig, and
// the true problem's vectors, rc rl ru rb): staging another instance of the
// same LP from it copies on the device.
struct DevSource {
  const int32_t *t_ptr = nullptr, *t_col = nullptr, *a_ptr = nullptr;
  const double *t_val = nullptr;  // unscaled CSC values
  const double *c = nullptr, *l = nullptr, *u = nullptr, *b = nullptr;
};

// Per-instance host state of the optimize() loop.
struct HostInst {
  Prepared P;
  InstDev dev{};
  void *arena = nullptr;
  size_t arena_bytes = 0;
  int64_t eval_iter = 1;    // value of the reference's `iteration` at the next evaluation
  int64_t acc_at_eval = 0;  // accepted steps at the last evaluation
  int64_t iter_start = 1;   // `iteration` of the last evaluation
  bool needs_eval = true;
  bool done = false;
  int32_t reason = 0;
  double start_time = 0.0, basic_time = 0.0;
  double last_trial = 1.0, pdist = 0.0, ddist = 0.0;
  pdlp_iteration_stats last_stats{};
  pdlp_result *res = nullptr;
  pdlp_stats_cb cb = nullptr;
  void *user = nullptr;
  double t_submit = 0.0;             // wall time the request entered the library
  // FP32 + refinement state
  QpCache qc_cur{};                  // norms of the LP the current solve terminates on
  double tol_cur = -1.0;             // tolerance of the current (inner) solve, -1 = user's
  int32_t ref_round = 0;             // correction LPs solved so far
  int32_t capped_rounds = 0;         // correction solves stopped by their iteration cap
  int32_t capped_seen = 0;           // capped_rounds at the last refinement test
  bool stalled = false;              // refinement stalled: finished for an FP64 solve
  bool refine_now = false;           // inner solve finished in this evaluation
  bool refined = false;              // (xk, yk) holds a refined point
  double dp = 1.0, dd = 1.0;         // scale factors of the current correction LP
  int64_t inner_cap = 0;             // iteration cap of the current correction solve (0 none)
  double gap_norm_cur = -1.0;        // gap-test normalisation of the current solve (-1: |p|+|d|)
  double best_err = INFINITY;        // KKT error of the best refined point (xb, yb)
  pdlp_iteration_stats best_stats{};
  double last_err = INFINITY;        // KKT error of (xk, yk)
  int64_t iters_total = 0, attempts_total = 0;
  double tolerance_override = -1.0;  // serving: per-request tolerance
  int64_t iteration_limit_override = -1;  // >= 0: tighter iteration limit for this instance
  double time_limit_override = -1.0;      // >= 0: tighter time limit (s) for this instance
  int64_t tag = -1;                  // serving: request id
  // Between Engine::stage and activation: the uploaded originals the device
  // preparation starts from (pool buffers on the FP32 path, the arena's own
  // arrays on the FP64 path) and the slice tables.
  bool staged = false;
  struct Stage {
    double *t_val = nullptr, *c = nullptr, *l = nullptr, *u = nullptr, *b = nullptr;
    SliceDesc *tab = nullptr;
    std::vector<void *> bufs;  // pool buffers to return
  } stage;
};

// Event-based timing of the iteration kernels (see pdlp_kernel_profile).
struct Profiler {
  bool enabled = false;
  bool timed = true;  // false: launches and bytes only, no events
  pdlp_kernel_profile acc{};
  std::vector<hipEvent_t> events;  // reusable pool
};

class Engine {
 public:
  Engine(hipStream_t s, DevicePool *pool, const pdlp_params &prm, Profiler *prof = nullptr);
  ~Engine();

  // Adds an instance (already prepared/rescaled on the host): stage + activation.
  int add(std::unique_ptr<HostInst> inst, std::string *err);
  // Copies an instance to the device (from the host, or from `src` on this
  // device) without rescaling or solving it; the next round, or
  // activate_staged, prepares it on the device and starts it.
  int stage(std::unique_ptr<HostInst> inst, std::string *err, const DevSource *src = nullptr);
  int activate_staged(std::string *err);
  // Runs all instances to termination.  Returns 0 or a negative HIP error.
  int run(std::string *err);
  // Runs one round: pending evaluations, then one chunk of attempts.  Returns
  // 0 while work remains, 1 when all instances are done, < 0 on error.
  int step_round(std::string *err);
  // The same in two halves, so that engines on different streams overlap:
  // begin_round evaluates and enqueues the chunk (same return values),
  // end_round waits for it.
  int begin_round(std::string *err);
  int end_round(std::string *err);
  bool all_done() const;
  // True when the chunk enqueued by begin_round has finished on the device.
  bool round_ready() const;
  size_t count() const { return inst_.size(); }
  HostInst &inst(size_t i) { return *inst_[i]; }
  void remove_done();  // drops finished instances (serving)
  // Ends instance q between rounds without writing its result (a staged
  // instance is dropped before its preparation).
  int cancel(size_t q, std::string *err);
  // Step attempts instance q has made in its current solve (as of the last
  // completed round).
  int64_t attempts(size_t q) const { return ctl_[q].attempts; }
  // Host staging buffers from the heap instead of pinned memory (an engine
  // that must not allocate through HIP); set before the first instance.
  void set_pageable_host(bool on) { pageable_host_ = on; }

 private:
  int transfer(HostInst &h, const DevSource *src, std::string *err);
  int prepare(HostInst &h, std::string *err);
  int activate(size_t q, std::string *err);
  int fill_slices(HostInst &h, std::string *err, const SliceDesc *dev_tab = nullptr);
  int constant_step(HostInst &h, double *eta, double *passes, std::string *err);
  int rebuild_tables(std::string *err);
  hipError_t push_ctl();
  int ensure_step_table(int64_t need, std::string *err);
  int evaluate(std::string *err);
  int launch_attempts(int64_t count, std::string *err);
  int complete_attempts(std::string *err);
  int finish_instance(size_t q, std::string *err);
  void host_eval_logic(size_t q, const double *out);
  int refine(std::string *err);
  void refine_logic(size_t q, const double *out);
  void reset_inner(size_t q, double omega0);

  hipError_t host_alloc(void **p, size_t bytes);
  void host_free(void *p);

  hipStream_t s_;
  DevicePool *pool_;
  Profiler *prof_ = nullptr;
  bool pageable_host_ = false;
  pdlp_params prm_;
  int prec_ = 0;
  // LDS reserved per work-group of the primal / dual slice kernels (occupancy caps)
  unsigned primal_lds_ = 0, dual_lds_ = 0;
  std::vector<std::unique_ptr<HostInst>> inst_;
  std::vector<Ctl> ctl_;  // host mirror (pinned copy below)
  Ctl *ctl_pinned_ = nullptr;
  double *out_pinned_ = nullptr;
  double *norm_pinned_ = nullptr;
  double *rout_pinned_ = nullptr;
  double *d_rout_ = nullptr;  // owned here (not a table: tables are rebuilt)
  size_t rout_cap_ = 0;
  size_t out_cap_ = 0;
  // device tables
  bool tables_dirty_ = true;
  int np_ = 0, nd_ = 0, nps_ = 0, nds_ = 0, npl_ = 0, ndl_ = 0, nlong_ = 0, nslots_ = 0;
  SliceDesc *d_pslices_ = nullptr, *d_dslices_ = nullptr;
  BlockDesc *d_pblocks_ = nullptr, *d_dblocks_ = nullptr;
  LongRow *d_longs_ = nullptr;
  int32_t *d_long_cnt_ = nullptr;
  double *d_long_part_ = nullptr;
  InstDev *d_insts_ = nullptr;
  Ctl *d_ctl_ = nullptr;
  double *d_ppart_ = nullptr, *d_dpart_ = nullptr, *d_epn_ = nullptr, *d_epm_ = nullptr,
         *d_epr_ = nullptr, *d_out_ = nullptr;
  double *d_red_ = nullptr, *d_grow_ = nullptr;
  int64_t tab_len_ = 0;
  std::vector<void *> table_allocs_;
  // the chunk in flight (launch_attempts -> complete_attempts)
  int64_t pend_count_ = 0;
  double pend_t0_ = 0.0;
  bool pend_prof_ = false;
  std::vector<double> pend_pbytes_, pend_dbytes_;
  std::vector<int64_t> pend_remaining_;
};

double now_sec();

}  // namespace pdlp
