// Group-mapped kernels for N = 11 bubbles per system (all harmonic counts and field types).
#include "cbc_group_kernel.h"

namespace cbc {
const void* group_pick_n11(int nk, int field) { return group_pick<11>(nk, field); }
}  // namespace cbc
