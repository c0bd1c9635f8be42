// Postsolve entry points (reference cpp/include/gpu_presolver/presolve/gpu_postsolve.hpp:11-27).
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

#include <vector>

namespace gpu_presolver::presolve {

// Structural-only primal restore (reference gpu_postsolve.hpp:11-14).
void postsolve_restore_structural_primal_gpu(
    double* x_org,
    const std::vector<StructuralL1PrimalRecoveryStep>& recoveries,
    const double* original_l = nullptr);

// Maps a reduced-space (x, y, z) back to the original space.  Any of the
// three inputs may be null.  The returned arrays are hipMalloc'ed and owned
// by the caller (reference gpu_postsolve.hpp:16-27).
GpuPostsolveResult postsolve_gpu(double* x_red,
                                 double* y_red,
                                 double* z_red,
                                 const PresolveRecordGpu& record,
                                 const double* original_l = nullptr);

GpuPostsolveResult postsolve_gpu(double* x_red,
                                 double* y_red,
                                 double* z_red,
                                 const PresolveRecordGpu& record,
                                 const LPInfoGpu* original_model,
                                 const double* original_l = nullptr);

}  // namespace gpu_presolver::presolve
