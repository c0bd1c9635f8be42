// Presolve entry points (reference cpp/include/gpu_presolver/presolve/gpu_presolve.hpp:9-40).
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

#include <cstdint>

namespace gpu_presolver::presolve {

// Result of one presolve call.  Move-only; owns reduced_lp when
// owns_reduced_lp is set (reference gpu_presolve.hpp:9-29).
struct GpuPresolveSummary {
  GpuPresolveSummary() = default;
  GpuPresolveSummary(const GpuPresolveSummary&) = delete;
  GpuPresolveSummary& operator=(const GpuPresolveSummary&) = delete;
  GpuPresolveSummary(GpuPresolveSummary&& other) noexcept;
  GpuPresolveSummary& operator=(GpuPresolveSummary&& other) noexcept;

  std::int32_t original_rows = 0;
  std::int32_t original_cols = 0;
  std::int32_t reduced_rows = 0;
  std::int32_t reduced_cols = 0;
  std::int32_t reduced_nnz = 0;
  std::int32_t iterations = 0;
  double elapsed_seconds = 0.0;
  double obj_constant_delta = 0.0;
  bool has_infeasible = false;
  bool has_unbounded = false;
  PresolveRecordGpu record;
  LPInfoGpu reduced_lp;
  bool owns_reduced_lp = false;
};

// True when nnz_after < ratio * nnz_before (reference gpu_presolve.hpp:32).
bool _has_good_nnz_progress(std::int32_t nnz_before, std::int32_t nnz_after, double ratio);

// Presolve keeping only the record; the reduced LP is released
// (reference gpu_presolve.hpp:34).
GpuPresolveSummary run_gpu_presolve_with_record(const LPInfoGpu& lp,
                                                const PresolveParams& params);

// Presolve keeping the reduced LP in device memory (reference gpu_presolve.hpp:37).
GpuPresolveSummary run_gpu_presolve_with_reduced_lp(const LPInfoGpu& lp,
                                                    const PresolveParams& params);

// Idempotent release of summary.reduced_lp (reference gpu_presolve.hpp:40).
void free_gpu_presolve_reduced_lp(GpuPresolveSummary& summary);

}  // namespace gpu_presolver::presolve
