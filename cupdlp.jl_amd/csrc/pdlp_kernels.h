// pdlp_kernels.h — host-callable launchers of the HIP kernels (pdlp_kernels.hip,
// pdlp_precond.hip).  Each returns a hipError_t value (0 = success).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "pdlp_device.h"

namespace pdlp {

// prec: 0 = FP64 iterates/matrix, 1 = FP32.
// Step kernels over the slice tables (one wave per slice): normal slices,
// then the chunk slices of long rows (their partials follow the normal
// slices' in part).
// unroll (4 or 8: loads in flight per lane) and waves (1, 2 or 4 per
// work-group) pick the variant of the normal-slice kernel; the variants give
// the same results.  lds_bytes (unused by the kernel) is reserved per
// work-group of the normal-slice kernel to cap how many share a CU
// (slice_lds_bytes).  Optional events e0 / e1 are recorded at the start of
// the first and the end of the last kernel's execution.
int launch_primal(hipStream_t s, int prec, int nslices, const SliceDesc *slices, int nlong,
                  const SliceDesc *lslices, const LongRow *longs, int32_t *long_cnt,
                  double *long_part, const InstDev *insts, const Ctl *ctl, double *part,
                  int unroll = 8, int waves = 4, unsigned lds_bytes = 0,
                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Default occupancy caps of the slice kernels (waves per SIMD), measured on
// MI355X: n = 1M LP FP64 k_dual 79 us at 3 vs 88 us uncapped, k_primal 91 us
// at 4-5 vs 94; 512-LP batch 60.1 vs 52.7 solves/s (FP64).
constexpr int kPrimalOccupancy = 5;
constexpr int kDualOccupancy = 3;
// LDS bytes per work-group of `waves` waves that hold the slice kernels to
// at most `occupancy` waves per SIMD on a device with lds_per_cu bytes of
// LDS per CU and lds_per_block per work-group (0 for occupancy <= 0).
unsigned slice_lds_bytes(int occupancy, int waves, int lds_per_cu, int lds_per_block);
int launch_dual(hipStream_t s, int prec, int nslices, const SliceDesc *slices, int nlong,
                const SliceDesc *lslices, const LongRow *longs, int32_t *long_cnt,
                double *long_part, const InstDev *insts, const Ctl *ctl, double *part,
                int unroll = 8, int waves = 4, unsigned lds_bytes = 0,
                hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Evaluation preparation over the row-block tables: pending weighted sums,
// and A^T y of the current y when it is not valid.
int launch_primal_flush(hipStream_t s, int prec, int nblocks, const BlockDesc *blocks,
                        const LongRow *longs, int32_t *long_cnt, double *long_part,
                        const InstDev *insts, const Ctl *ctl);
int launch_dual_flush(hipStream_t s, int prec, int nblocks, const BlockDesc *blocks,
                      const InstDev *insts, const Ctl *ctl);
// Sliced copy (scol, sval) of the CSR matrix (ptr, col, val) for the given
// instance-local slices.
int launch_slice_fill(hipStream_t s, int prec, int nslices, const SliceDesc *slices,
                      const int32_t *ptr, const int32_t *col, const void *val, int32_t *scol,
                      void *sval);
int launch_step(hipStream_t s, int ninst, const InstDev *insts, Ctl *ctl, const double *ppart,
                const double *dpart, const double *red_tab, const double *grow_tab,
                int64_t tab_len, double red_exp, double grow_exp, int constant,
                int64_t debug_attempt = 0);
int launch_eval(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl, double *epn,
                double *epm, double *epr, int ninst, double *out);
int launch_restart(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                   const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl);

// Refinement (FP32 path).  phase 1: update (xk, yk) from the inner solution,
// FP64 residual products and statistics (out: kRefOutStride doubles per
// instance, instances with ref_flag == 1); phase 2: set up the next
// correction LP (ref_flag == 2).  k_eval_finish checks eval_flag, which the
// caller keeps set for the refined instances.
constexpr int kRefN = 9;   // n-side statistics of the refined point
constexpr int kRefM = 9;   // m-side statistics
constexpr int kRefOutStride = kRefN + kRefM;
int launch_refine(hipStream_t s, int prec, int phase, int np, const BlockDesc *pblocks, int nd,
                  const BlockDesc *dblocks, const LongRow *longs, int32_t *long_cnt,
                  double *long_part, const InstDev *insts, const Ctl *ctl, double *epn,
                  double *epm, int ninst, double *out);

}  // namespace pdlp
