// Fused persistent kernels over the bins N = 8, 4, 3, 2 with 2 drive harmonic(s), all field types.
#include "cbc_fused_kernel.h"

namespace cbc {
const void* fused_pick_8432_k2(int field) { return fused_pick<2, 8, 4, 3, 2>(field); }
}  // namespace cbc
