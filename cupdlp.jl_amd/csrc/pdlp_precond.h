// pdlp_precond.h — device-side problem preparation (see pdlp_precond.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "pdlp_device.h"

namespace pdlp {

struct PrecondArgs {
  int64_t n = 0, m = 0, nnz = 0;
  const int32_t *t_ptr = nullptr, *t_col = nullptr;  // CSC of A (= CSR of A^T)
  double *t_val = nullptr;                           // in: original values; out: scaled
  const int32_t *a_ptr = nullptr;                    // CSR row pointers (from the host)
  int32_t *a_col = nullptr, *cursor = nullptr;       // out: CSR columns; scratch (m, zeroed)
  double *a_val = nullptr;                           // out: scaled CSR values
  double *c = nullptr, *l = nullptr, *u = nullptr, *b = nullptr;  // in: original; out: scaled
  double *D = nullptr, *E = nullptr;                 // in: ones; out: cumulative scaling
  double *r = nullptr, *v = nullptr;                 // scratch (m), (n)
  double *norm_part = nullptr;                       // scratch, 5 * 256 doubles
  double *norm_host = nullptr;                       // pinned host, 5 * 256 doubles
  // instance-local slice tables of A^T (ps) and A (ds) on the device; the
  // row kernels walk them
  const SliceDesc *ps = nullptr, *ds = nullptr;
  int nps = 0, nds = 0;
  int64_t ruiz_iterations = 10;
  bool has_pc = true;
  double pc_alpha = 1.0;
  // Called after the CSR scatter and before any scaling (e.g. to keep the
  // unscaled CSR values); may be null.
  hipError_t (*after_scatter)(hipStream_t, void *) = nullptr;
  void *after_scatter_ctx = nullptr;
};

struct PrecondOut {
  double max_abs = 0.0;  // norm(scaled A, Inf), elementwise
  double norm_b = 0.0;   // ||scaled b||_2
  double norm_c = 0.0;   // ||scaled c||_2
};

hipError_t precond_run(hipStream_t s, const PrecondArgs &a, PrecondOut *out);

// estimate_maximum_singular_value (src/primal_dual_hybrid_gradient_gpu.jl:150-172)
// on the device: `iterations` rounds of x <- A^T (A (x / ||x||)) from the
// start vector in x, then sqrt(x'A^T A x / ||x||^2), every sum exactly
// rounded.  x (n), t (m), z (n) device scratch; part 512 doubles; part_host
// pinned, 512 doubles.
struct PowerArgs {
  int64_t n = 0, m = 0;
  const SliceDesc *ps = nullptr, *ds = nullptr;
  int nps = 0, nds = 0;
  const int32_t *t_ptr = nullptr, *t_col = nullptr, *a_ptr = nullptr, *a_col = nullptr;
  const double *t_val = nullptr, *a_val = nullptr;
  double *x = nullptr, *t = nullptr, *z = nullptr, *part = nullptr, *part_host = nullptr;
  int64_t iterations = 0;
};
hipError_t power_method(hipStream_t s, const PowerArgs &a, double *sv);
hipError_t to_float(hipStream_t s, int64_t len, const double *src, float *dst);

}  // namespace pdlp
