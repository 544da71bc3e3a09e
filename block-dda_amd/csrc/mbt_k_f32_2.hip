// mbt_k_f32_2.hip -- FP32 kernels of the planned passes for group 2 of the plan table
// (explicit instantiations of the per-plan launch functions of mbt_plans.hpp;
// the table's groups compile as separate units, in parallel).
#define BDDA_PLAN_INSTANTIATE
#include "mbt_plans.hpp"

namespace bdda {
#define BDDA_PLAN_DEF(p, ...) BDDA_PLAN_INST(, float, PlanOf##p)
BDDA_PLAN_TABLE_2(BDDA_PLAN_DEF)
#undef BDDA_PLAN_DEF
BDDA_PLAN_INST_X(, float, PlanX384F32, 3)
BDDA_PLAN_INST_X(, float, PlanX384F32, 6)
}  // namespace bdda
