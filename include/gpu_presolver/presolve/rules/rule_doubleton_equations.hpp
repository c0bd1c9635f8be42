// Rule operator API, see reference cpp/include/gpu_presolver/presolve/rules/rule_doubleton_equations.hpp.
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

void apply_rule_doubleton_equations(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                    const PresolveStatsGpu& stats, const PresolveParams& pparams);

}  // namespace gpu_presolver::presolve
