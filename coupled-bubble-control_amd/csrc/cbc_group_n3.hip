// Group-mapped kernels for N = 3 bubbles per system (all harmonic counts and field types).
#include "cbc_group_kernel.h"

namespace cbc {
const void* group_pick_n3(int nk, int field) { return group_pick<3>(nk, field); }
}  // namespace cbc
