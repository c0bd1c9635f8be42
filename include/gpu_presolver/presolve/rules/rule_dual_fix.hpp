// Rule operator API, see reference cpp/include/gpu_presolver/presolve/rules/rule_dual_fix.hpp.
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

void apply_rule_dual_fix(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                         const PresolveStatsGpu& stats, const PresolveParams& pparams,
                         const std::uint8_t* candidate_col_mask = nullptr);

}  // namespace gpu_presolver::presolve
