// mbt_k_f64_3.hip -- FP64 kernels of the planned passes for group 3 of the plan table
// (explicit instantiations of the per-plan launch functions of mbt_plans.hpp;
// the table's groups compile as separate units, in parallel).
#define BDDA_PLAN_INSTANTIATE
#include "mbt_plans.hpp"

namespace bdda {
#define BDDA_PLAN_DEF(p, ...) BDDA_PLAN_INST(, double, PlanOf##p)
BDDA_PLAN_TABLE_3(BDDA_PLAN_DEF)
#undef BDDA_PLAN_DEF
}  // namespace bdda
