// Group-mapped kernels for N = 16 bubbles per system (all harmonic counts and field types).
#include "cbc_group_kernel.h"

namespace cbc {
const void* group_pick_n16(int nk, int field) { return group_pick<16>(nk, field); }
}  // namespace cbc
