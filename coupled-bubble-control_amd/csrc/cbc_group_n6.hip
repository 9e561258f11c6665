// Group-mapped kernels for N = 6 bubbles per system (all harmonic counts and field types).
#include "cbc_group_kernel.h"

namespace cbc {
const void* group_pick_n6(int nk, int field) { return group_pick<6>(nk, field); }
}  // namespace cbc
