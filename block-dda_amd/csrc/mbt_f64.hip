// mbt_f64.hip -- FP64 instantiation of the MBT pass sequence (mbt_passes.hpp).
#include "mbt_passes.hpp"

namespace bdda {
template void mbt_apply_precision<double>(MbtOperator*, const cplx<double>*, int64_t, cplx<double>*, int64_t, int,
                                        const cplx<double>*, const cplx<double>*, const int*);
}  // namespace bdda
