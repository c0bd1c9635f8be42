// Rule operator API, see reference cpp/include/gpu_presolver/presolve/rules/rule_duplicate_columns.hpp.
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

void apply_rule_duplicate_columns(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                  const PresolveStatsGpu& stats, const PresolveParams& pparams);

}  // namespace gpu_presolver::presolve
