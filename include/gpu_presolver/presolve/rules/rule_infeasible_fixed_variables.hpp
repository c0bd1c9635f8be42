// Rule operator API, see reference cpp/include/gpu_presolver/presolve/rules/rule_infeasible_fixed_variables.hpp.
#pragma once

#include "gpu_presolver/presolve/presolve_structs.hpp"

namespace gpu_presolver::presolve {

// Fixes columns with |u - l| <= bound_tol at their midpoint.
void apply_rule_infeasible_fixed_variables(PresolvePlanGpu& plan, const LPInfoGpu& lp,
                                           const PresolveParams& pparams);

}  // namespace gpu_presolver::presolve
