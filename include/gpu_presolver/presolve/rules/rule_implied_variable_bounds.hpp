// Rule operator API, see reference cpp/include/gpu_presolver/presolve/rules/rule_implied_variable_bounds.hpp.
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

void apply_rule_implied_variable_bounds(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                        const PresolveStatsGpu& stats,
                                        const PresolveParams& pparams);

// Releases scratch retained between implied-bound rounds of one presolve.
void release_implied_variable_bounds_workspace() noexcept;

}  // namespace gpu_presolver::presolve
