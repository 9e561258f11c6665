// Group-mapped kernels for N = 9 bubbles per system (all harmonic counts and field types).
#include "cbc_group_kernel.h"

namespace cbc {
const void* group_pick_n9(int nk, int field) { return group_pick<9>(nk, field); }
}  // namespace cbc
