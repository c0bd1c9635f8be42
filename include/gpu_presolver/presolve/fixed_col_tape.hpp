// FixedCol tape packing (reference cpp/include/gpu_presolver/presolve/fixed_col_tape.hpp:13-21).
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

#include <cstdint>

namespace gpu_presolver::presolve {

// Appends one FixedCol record {col; v, c, (row, a) over live rows} per
// selected column (fixed_mask set, excluded_mask clear or null) in ascending
// column order, and returns sum_j c_j v_j accumulated in long double.  `tape`
// may be null when only the objective delta is wanted.
double append_compact_fixed_col_tape_from_device(PostsolveTape* tape,
                                                 const std::uint8_t* fixed_mask,
                                                 const std::uint8_t* excluded_mask,
                                                 const double* fixed_val,
                                                 const std::uint8_t* keep_row,
                                                 const double* c,
                                                 const DeviceCsrMatrix& AT,
                                                 std::int32_t n,
                                                 const char* context);

}  // namespace gpu_presolver::presolve
