// mbt_f32.hip -- FP32 instantiation of the MBT pass sequence (mbt_passes.hpp).
#include "mbt_passes.hpp"

namespace bdda {
template void mbt_apply_precision<float>(MbtOperator*, const cplx<float>*, int64_t, cplx<float>*, int64_t, int,
                                        const cplx<float>*, const cplx<float>*, const int*);
}  // namespace bdda
