// pdlp_device.h — device-side data layout shared by the HIP kernels and the
// host engine.  No HIP types appear here, so the host C++ can include it.
//
// Layout in HBM (per LP "instance"; a single solve is a batch of one):
//   A   (m x n) CSR : a_ptr[m+1] int32, a_col[nnz] int32, a_val[nnz] T
//   A^T (n x m) CSR : t_ptr[n+1] int32, t_col[nnz] int32, t_val[nnz] T
//   sliced A, A^T   : as_col/ts_col int32, as_val/ts_val T (see SliceDesc)
//   scaled data     : c, l, u (n) and b (m), type T
//   original data   : oc, ol, ou (n), ob (m), D (n), E (m), double
//   iterates        : x[2], aty, dx (n); y[2], ax[2] (m)          (T)
//   weighted sums   : sx, saty (n); sy, sax (m)                      (T)
//   evaluation      : avg_x, avg_g, gbuf, lx, lg (n); avg_y, avg_ax, ly, lax (m)
// Rows of A and of A^T are cut into "row blocks" of at most kBlockNnz
// nonzeros (and kBlockRows rows); a row longer than kBlockNnz is split into
// chunks whose partial sums are combined in chunk order by the last chunk.
// Blocks never straddle instances, so an instance's arithmetic does not
// depend on what else is packed beside it.
#pragma once
#include <stdint.h>

namespace pdlp {

constexpr int kThreads = 256;    // work-group size of every kernel (4 waves)
constexpr int kBlockNnz = 1024;  // nonzeros staged in LDS per row block
constexpr int kBlockRows = 512;  // rows per row block
// A row longer than kBlockNnz gets a work-group of its own; a row longer than
// kLongChunk is split into kLongChunk-entry chunks combined by the last chunk
// (the only cross-work-group handoff in the iteration kernels).
constexpr int kLongChunk = 65536;

// Evaluation accumulators (see eval kernel): n-side and m-side.
constexpr int kEvalN = 25;
constexpr int kEvalM = 22;
constexpr int kEvalOut = kEvalN + kEvalM;  // index of the primal ray objective
constexpr int kEvalStride = kEvalOut + 1;   // doubles per instance in the output

// Sliced layout of the step kernels (one wave per slice, no LDS, no
// barriers).  A slice covers `nrows` consecutive rows with 2^lg lanes per
// row; entry k of the slice's row g sits at slot off + (k >> lg) * 64 +
// (g << lg) + (k & (2^lg - 1)), so each load instruction of a wave reads 64
// consecutive slots.  Unused slots hold column -1.  Rows of up to
// 64 * kSellWidth entries get the fewest lanes (a power of two) for which a
// lane sums at most kSellWidth entries; longer rows get a whole wave, and a
// row longer than kSellChunk entries is cut into chunk slices (nrows = 0)
// whose partial sums the last chunk combines in chunk order.
constexpr int kSellWidth = 64;
constexpr int kSellChunk = 16384;

struct SliceDesc {
  int32_t inst;   // instance index
  int32_t row0;   // first row (chunk slices: the long row)
  int32_t off;    // first slot in the instance's sliced arrays
  int32_t width;  // slots per lane
  int16_t lg;     // log2 of the lanes per row (chunk slices: 6)
  int16_t nrows;  // rows in the slice; 0 = chunk slice; < 0 = no work
  int32_t lid;    // chunk slices: long-row id
  int32_t chunk;  // chunk slices: chunk index
  int32_t part;   // index of the slice's partial sums (instance-contiguous)
};

struct BlockDesc {
  int32_t inst;   // instance index
  int32_t row0;   // first row (or the long row)
  int32_t nrows;  // > 0: rows [row0, row0+nrows); < 0: chunk (-nrows-1) of long row row0
  int32_t aux;    // long-row id for chunks
};

struct LongRow {
  int32_t slot;     // first partial-sum slot
  int32_t nchunks;  // number of chunks
};

struct InstDev {
  int32_t n, m, neq, pad0;
  int32_t pb0, pb1;  // block range in the A^T (primal) table
  int32_t db0, db1;  // block range in the A (dual) table
  const int32_t *a_ptr, *a_col;
  const void *a_val;
  const int32_t *t_ptr, *t_col;
  const void *t_val;
  const void *c, *l, *u, *b;  // scaled (T)
  const double *oc, *ol, *ou, *ob, *D, *E;
  void *x0, *x1, *y0, *y1, *ax0, *ax1, *aty, *dx;
  void *sx, *saty, *sy, *sax;
  void *avg_x, *avg_g, *gbuf, *lx, *lg;
  void *avg_y, *avg_ax, *ly, *lax;
  const double *a_orig, *t_orig;  // FP32 path: unscaled FP64 values (CSR, CSC order)
  // sliced copies of A (dual step) and A^T (primal step), and the ranges of
  // this instance in the slice tables (step partials are per slice)
  const int32_t *as_col, *ts_col;
  const void *as_val, *ts_val;
  int32_t ps0, ps1, ds0, ds1;  // normal slices
  int32_t pl0, pl1, dl0, dl1;  // chunk slices (indices into the same partial arrays)
  // FP32 + refinement: dual lower bounds of ">=" rows (scaled T, original FP64;
  // null = 0), the refined FP64 point, its products, and the true problem data
  void *ylo;
  const double *oylo;
  double *xk, *yk, *rax, *raty;
  const double *rc, *rl, *ru, *rb;
  double *xb, *yb;  // the best refined point so far
};

// Per-instance control block.  Written by the step kernel and by the host
// between chunks (the stream is idle then).
struct Ctl {
  double eta;        // step size of the next attempt (take_step!'s local step_size)
  double eta_entry;  // solver_state.step_size (weight of the accepted point)
  double omega;      // primal weight
  double w_pending;  // weight of an accepted point whose sums are not yet added
  double wsum;       // sum of weights
  double kkt_passes; // cumulative_kkt_passes
  double last_interaction, last_movement;
  int64_t count;     // sum_primal_solutions_count
  int64_t attempts;  // total_number_iterations (attempts)
  int64_t accepted;  // completed take_step! calls
  int64_t target;    // stop when accepted reaches this
  int32_t cur;       // which of x0/x1 (y0/y1, ax0/ax1) is current
  int32_t pending;   // sums of the last accepted point still to add
  int32_t aty_valid; // aty holds A^T y for the current y
  int32_t active;    // attempts run for this instance
  int32_t numerical_error;
  int32_t eval_flag;      // instance is being evaluated
  int32_t use_current;    // evaluation averages := current point
  int32_t restart_action; // 0 none, 1 reset sums to current, 2 restart to average
  int32_t ref_flag;       // refinement kernels act on this instance
  int32_t ref_keep;       // the setup copies (xk, yk) to the best point (xb, yb)
  double ref_dp, ref_dd;  // refinement scale factors of the primal and dual correction
};

}  // namespace pdlp
