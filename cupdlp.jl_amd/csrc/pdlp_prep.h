// pdlp_prep.h — host-side problem preparation: validation, the cached norms
// of the original problem, diagonal rescaling and the CSR layouts the
// kernels read.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "pdlp_device.h"
#include "pdlp_hip.h"

namespace pdlp {

struct QpCache {  // CachedQuadraticProgramInfo, src/termination.jl:123-137
  double l_inf_c, l_inf_b, l2_c, l2_b;
};

// The slices of one matrix layout (see SliceDesc): descriptors with inst = 0
// and instance-local long-row ids, the chunk count of each long row, and the
// number of slots of the sliced arrays.
struct SliceTable {
  std::vector<SliceDesc> sl;
  std::vector<int32_t> nchunks;
  int64_t slots = 0;
};

// Cuts rows [0, R) with row pointers ptr into slices: a slice takes
// consecutive rows while each fits 2^lg * kSellWidth entries for the lane
// count 2^lg chosen by its first row (and at most 64 >> lg rows); a row of
// more than 64 * kSellWidth entries is a slice of its own, one of more than
// kSellChunk entries becomes chunk slices.
void build_slices(const int32_t *ptr, int64_t R, SliceTable *out, int lane_budget = 0);

// An LP after rescale_problem, in the device layouts (0-based, int32).
struct Prepared {
  int64_t n = 0, m = 0, neq = 0, nnz = 0;
  double c0 = 0.0;
  // A^T as CSR = A as CSC
  std::vector<int32_t> t_ptr, t_col;
  std::vector<double> t_val;
  // A as CSR
  std::vector<int32_t> a_ptr, a_col;
  std::vector<double> a_val;
  std::vector<int64_t> csr_from_csc;  // a_val[p] = t_val[csr_from_csc[p]]
  // scaled vectors and the scaling (E rows, D columns)
  std::vector<double> c, l, u, b, E, D;
  // original vectors
  std::vector<double> oc, ol, ou, ob;
  QpCache qc{};
  double max_abs_scaled = 0.0;  // norm(scaled A, Inf) (elementwise)
  std::vector<double> t_orig_host;  // unscaled CSC values (host-prepared FP32 path)
  // device numbering -> the caller's: column j here is column col_perm[j] of
  // the input, row i is row row_perm[i]
  std::vector<int32_t> col_perm, row_perm;
  SliceTable ps, ds;            // slices of A^T (primal step) and A (dual step)
  bool device_prep = false;     // rescaling and the CSR of A are built on the device
  double norm_b_scaled = 0.0, norm_c_scaled = 0.0;  // set by the device preparation
};

// validate (src/preprocess.jl:8-74) plus index checks of the C view.
// Returns 0 or a positive code with *err set.
int validate_lp(const pdlp_lp *lp, std::string *err);

// Loads the original problem and its cached norms.  With build_csr the CSR
// of A is built on the host; otherwise only its row pointers (the device
// builds the rest).  Rescaling is applied by rescale_host or on the device.
int load_problem(const pdlp_lp *lp, Prepared *P, std::string *err, bool build_csr = true,
                 int lane_budget = 0);

// rescale_problem (src/preprocess.jl:619-675) on the host: Ruiz-inf,
// optional L2, optional Pock-Chambolle.  Updates t_val/a_val and the vectors.
void rescale_host(const pdlp_params *prm, Prepared *P);

}  // namespace pdlp
