// Dispatch of the group-mapped kernels (cbc_group_kernel.h); the kernels are
// instantiated per bubble count (2..16) in cbc_group_n*.hip so they compile in parallel.
#include "cbc_kernels.h"

namespace cbc {

const void* group_pick_n2(int nk, int field);
const void* group_pick_n3(int nk, int field);
const void* group_pick_n4(int nk, int field);
const void* group_pick_n5(int nk, int field);
const void* group_pick_n6(int nk, int field);
const void* group_pick_n7(int nk, int field);
const void* group_pick_n8(int nk, int field);
const void* group_pick_n9(int nk, int field);
const void* group_pick_n10(int nk, int field);
const void* group_pick_n11(int nk, int field);
const void* group_pick_n12(int nk, int field);
const void* group_pick_n13(int nk, int field);
const void* group_pick_n14(int nk, int field);
const void* group_pick_n15(int nk, int field);
const void* group_pick_n16(int nk, int field);

static const void* group_pick_any(int ups, int nk, int field)
{
    switch (ups) {
    case 2: return group_pick_n2(nk, field);
    case 3: return group_pick_n3(nk, field);
    case 4: return group_pick_n4(nk, field);
    case 5: return group_pick_n5(nk, field);
    case 6: return group_pick_n6(nk, field);
    case 7: return group_pick_n7(nk, field);
    case 8: return group_pick_n8(nk, field);
    case 9: return group_pick_n9(nk, field);
    case 10: return group_pick_n10(nk, field);
    case 11: return group_pick_n11(nk, field);
    case 12: return group_pick_n12(nk, field);
    case 13: return group_pick_n13(nk, field);
    case 14: return group_pick_n14(nk, field);
    case 15: return group_pick_n15(nk, field);
    case 16: return group_pick_n16(nk, field);
    default: return nullptr;
    }
}

bool group_supported(int ups, int nk, int field) { return group_pick_any(ups, nk, field) != nullptr; }

int group_systems_per_block(int ups) { return 64 / group_lanes(ups); }

hipError_t group_launch(const LaunchArgs& a, hipStream_t stream)
{
    const void* fn = group_pick_any(a.ups, a.nk, a.field);
    if (!fn) return hipErrorInvalidValue;
    const int spb = group_systems_per_block(a.ups);
    const int grid = (a.ns + spb - 1) / spb;
    void* args[] = {const_cast<LaunchArgs*>(&a)};
    return hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, stream);
}

hipError_t group_launch_persistent(const LaunchArgs& a, int waves, hipStream_t stream)
{
    const void* fn = group_pick_any(a.ups, a.nk, a.field);
    if (!fn || !(a.queue || a.ring) || !a.queue_head || waves <= 0) return hipErrorInvalidValue;
    void* args[] = {const_cast<LaunchArgs*>(&a)};
    return hipLaunchKernel(fn, dim3(waves), dim3(64), args, 0, stream);
}

const void* fused_pick_8432_k1(int field);
const void* fused_pick_8432_k2(int field);
const void* fused_pick_16_2_k1(int field);
const void* fused_pick_16_2_k2(int field);

// Bubble-count lists with a fused kernel (largest first, 0 = unused position):
// the batched uneven ensemble (N = 2, 3, 4, 8) and the large-coupling set mixed
// with N = 2 systems.
struct FusedList {
    int n[kFusedMaxBins];
    const void* (*pick[2])(int field);  // by harmonic count 1, 2
};
static const FusedList kFusedLists[] = {
    {{8, 4, 3, 2}, {fused_pick_8432_k1, fused_pick_8432_k2}},
    {{16, 2, 0, 0}, {fused_pick_16_2_k1, fused_pick_16_2_k2}},
};

const void* fused_pick(const int* ups, int nb, int nk, int field, int* slot)
{
    if (nb < 1 || nb > kFusedMaxBins || nk < 1 || nk > 2) return nullptr;
    for (const FusedList& l : kFusedLists) {
        bool covered = true;
        for (int b = 0; b < nb && covered; ++b) {
            slot[b] = -1;
            for (int p = 0; p < kFusedMaxBins; ++p)
                if (l.n[p] == ups[b]) slot[b] = p;
            covered = slot[b] >= 0;
        }
        if (covered) return l.pick[nk - 1](field);
    }
    return nullptr;
}

hipError_t fused_launch(const void* fn, const FusedArgs& f, int waves, hipStream_t stream)
{
    if (!fn || waves <= 0) return hipErrorInvalidValue;
    void* args[] = {const_cast<FusedArgs*>(&f)};
    return hipLaunchKernel(fn, dim3(waves), dim3(64), args, 0, stream);
}

static hipError_t attr_of(const void* fn, KernelAttr* out)
{
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, fn);
    if (e != hipSuccess) return e;
    out->num_regs = fa.numRegs;
    out->local_bytes = (int32_t)fa.localSizeBytes;
    out->max_threads = fa.maxThreadsPerBlock;
    out->lds_bytes = (int32_t)fa.sharedSizeBytes;
    return hipSuccess;
}

hipError_t group_attr(int ups, int nk, int field, KernelAttr* out)
{
    const void* fn = group_pick_any(ups, nk, field);
    if (!fn) return hipErrorInvalidValue;
    return attr_of(fn, out);
}

hipError_t fused_attr(const void* fn, KernelAttr* out)
{
    if (!fn) return hipErrorInvalidValue;
    return attr_of(fn, out);
}

}  // namespace cbc
