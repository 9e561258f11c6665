// Fused persistent kernels over the bins N = 16, 2 with 2 drive harmonic(s), all field types.
#include "cbc_fused_kernel.h"

namespace cbc {
const void* fused_pick_16_2_k2(int field) { return fused_pick<2, 16, 2>(field); }
}  // namespace cbc
