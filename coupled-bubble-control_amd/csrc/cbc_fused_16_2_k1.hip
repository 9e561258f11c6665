// Fused persistent kernels over the bins N = 16, 2 with 1 drive harmonic(s), all field types.
#include "cbc_fused_kernel.h"

namespace cbc {
const void* fused_pick_16_2_k1(int field) { return fused_pick<1, 16, 2>(field); }
}  // namespace cbc
