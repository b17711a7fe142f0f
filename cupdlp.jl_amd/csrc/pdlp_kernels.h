// pdlp_kernels.h — host-callable launchers of the HIP kernels (pdlp_kernels.hip,
// pdlp_precond.hip).  Each returns a hipError_t value (0 = success).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "pdlp_device.h"

namespace pdlp {

// prec: 0 = FP64 iterates/matrix, 1 = FP32.
int launch_primal(hipStream_t s, int prec, bool flush, int nblocks, const BlockDesc *blocks,
                  const LongRow *longs, int32_t *long_cnt, double *long_part,
                  const InstDev *insts, const Ctl *ctl, double *part);
int launch_dual(hipStream_t s, int prec, bool flush, int nblocks, const BlockDesc *blocks,
                const LongRow *longs, int32_t *long_cnt, double *long_part, const InstDev *insts,
                const Ctl *ctl, double *part);
int launch_step(hipStream_t s, int ninst, const InstDev *insts, Ctl *ctl, const double *ppart,
                const double *dpart, const double *red_tab, const double *grow_tab,
                int64_t tab_len, double red_exp, double grow_exp);
int launch_eval(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl, double *epn,
                double *epm, double *epr, int ninst, double *out);
int launch_restart(hipStream_t s, int prec, int np, const BlockDesc *pblocks, int nd,
                   const BlockDesc *dblocks, const InstDev *insts, const Ctl *ctl);

}  // namespace pdlp
