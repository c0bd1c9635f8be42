// Shared activity primitives (reference cpp/include/gpu_presolver/presolve/gpu_presolve_kernels.hpp:7-16).
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

// Per row: sum of the finite min/max terms a*l or a*u over entries with
// |a| > zero_tol, plus the counts of -inf min terms and +inf max terms.
void compute_row_activity_summary(double* row_min_fin,
                                  double* row_max_fin,
                                  std::int32_t* row_min_neg_inf_count,
                                  std::int32_t* row_max_pos_inf_count,
                                  const DeviceCsrMatrix& A_csr,
                                  const double* l,
                                  const double* u,
                                  double zero_tol);

// max |a| over each row of AT, i.e. over each column of A.
void compute_col_max_abs(double* col_max_abs, const DeviceCsrMatrix& AT_csr);

}  // namespace gpu_presolver::presolve
