// cbc_fused_kernel.h -- one persistent launch over several bubble-count bins.
// Included by the translation units cbc_fused_*.hip, one per (bubble-count
// list, harmonic count).
//
// Every wave of the grid runs the group-mapped persistent stepper (group_run,
// cbc_group_kernel.h) on each bin of the list in turn, largest bubble count
// first: it takes systems from that bin's work queue until the queue is
// drained and its own groups have finished, then moves on to the next bin.
// The bins with the longest attempt chains start first, and waves that find a
// bin drained go on to the smaller ones, so the bins share the device in one
// dispatch without host synchronisation, side streams or the hardware queue
// those would need.  A bin runs exactly the code and arithmetic of its own
// group kernel, so its results are bit for bit those of a per-bin launch.
#pragma once

#include "cbc_group_kernel.h"

namespace cbc {

template <int... Ns>
constexpr size_t fused_lds_bytes()
{
    size_t m = 0;
    ((m = sizeof(GroupLds<Ns>) * GroupShape<Ns>::S > m ? sizeof(GroupLds<Ns>) * GroupShape<Ns>::S : m), ...);
    return m;
}

template <int NK, int FIELD, int B, int N, int... Rest>
__device__ __forceinline__ void fused_bins(const FusedArgs& f, unsigned char* pool)
{
    if (f.bin[B].queue_len > 0) {
        group_run<N, NK, FIELD>(f.bin[B], 0, reinterpret_cast<GroupLds<N>*>(pool));
        __syncthreads();  // the next bin lays its group slots over the same LDS
    }
    if constexpr (sizeof...(Rest) > 0) fused_bins<NK, FIELD, B + 1, Rest...>(f, pool);
}

// Ns: the bubble counts of the bins, largest first; bin b's arguments are f.bin[b]
// (queue_len == 0: no such bin in this launch).
template <int NK, int FIELD, int... Ns>
__global__ void __launch_bounds__(64) group_fused_kernel(const FusedArgs f)
{
    static_assert(sizeof...(Ns) <= kFusedMaxBins, "too many bins for FusedArgs");
    __shared__ __align__(16) unsigned char pool[fused_lds_bytes<Ns...>()];
    fused_bins<NK, FIELD, 0, Ns...>(f, pool);
}

template <int NK, int... Ns>
const void* fused_pick(int field)
{
    if (field == kFieldConst) return reinterpret_cast<const void*>(&group_fused_kernel<NK, kFieldConst, Ns...>);
    if (field == kFieldSwN) return reinterpret_cast<const void*>(&group_fused_kernel<NK, kFieldSwN, Ns...>);
    if (field == kFieldSwA) return reinterpret_cast<const void*>(&group_fused_kernel<NK, kFieldSwA, Ns...>);
    return nullptr;
}

}  // namespace cbc
