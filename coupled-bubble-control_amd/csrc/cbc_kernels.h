// cbc_kernels.h -- launch interface between the C ABI (cbc_api.hip) and the
// kernel translation units.  Host-only types; no HIP device code here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cbc_ring.h"

namespace cbc {

// Arguments of one ensemble launch in the reference buffer layout.
struct LaunchArgs {
    int32_t ns, tile, ups, ndo;
    int32_t nk, field;
    int64_t kernel_steps;      // attempts per system in this launch
    double atol, rtol, min_step, max_step, etol;
    double col_factor;         // 1 + collision threshold
    int event_exact;           // CBC_EVENT_EXACT
    int resume;                // a later launch of one solve call: systems whose flag is
                               // negative (a failure stop) stay stopped, untouched

    double* t;
    const double* t_end;
    const double* t_begin;
    double* h;
    double* state;
    const double* up;          // unit_params (13, NS, TILE)
    const double* sp;          // system_params (5, NS)
    const double* cp;          // control_params (4*NK, NS)
    const double* cm;          // coupling (3, UPS, NS, TILE)
    int32_t* dense_index;
    double* dense_time;
    double* dense_state;
    double* event;
    int32_t* flags;
    int32_t* total_steps;      // may be null
    int32_t* rejected_steps;   // may be null

    // period windows (period <= 0: off)
    double period;
    int32_t first_window;
    int32_t n_windows;
    double* max_r;
    double* mean_x;

    uint32_t* active_count;    // +1 per system left ACTIVE (may be null)

    // persistent mode (queue != null): lanes / lane groups take system indices
    // from queue[0..queue_len) through the atomic head counter and refill when
    // at least `refill` of them are free; kernel_steps is then the attempt
    // budget per system.
    const int32_t* queue;
    int32_t queue_len;
    int32_t refill;
    uint32_t* queue_head;

    // serving mode (ring != null, implies persistent): work items are taken from
    // a host-mapped ring that the host keeps appending to; queue_head counts the
    // items claimed.  See ServeItem.
    const struct ServeItem* ring;
    const uint32_t* ring_tail;   // host-mapped: number of items published
    const uint32_t* ring_stop;   // host-mapped: nonzero = exit once idle
    uint32_t ring_mask;          // ring capacity - 1 (capacity a power of two)
    int32_t poll_interval;       // attempts between ring polls of a wave with free lanes
                                 // (up to 8x that after polls that claim nothing)
    uint64_t idle_ticks;         // wall-clock ticks an idle wave waits for work before exiting
    uint32_t* done_seq;          // host-mapped (NS): seq of the last item finished per system
    double* out_state;           // host-mapped (4, NS, TILE) final state copy (may be null)
    double* out_misc;            // host-mapped (NS, 4): t, h, flag, attempts of the last item
    // host-mapped staging mirrors (reference layouts) read by items with kServeReset
    const double* stage_state;   // (4, NS, TILE)
    const double* stage_up;      // (13, NS, TILE)
    const double* stage_cm;      // (3, UPS, NS, TILE)
};

// ServeItem and its option bits: cbc_ring.h

// Fused persistent launch over up to kFusedMaxBins bins (one per bubble count),
// each a queued launch of its own plan: bin b of the kernel's bubble-count list
// takes its arguments from bin[b] (queue_len == 0: the bin is absent).
constexpr int kFusedMaxBins = 4;
struct FusedArgs {
    LaunchArgs bin[kFusedMaxBins];
};

struct KernelAttr {
    int32_t num_regs;
    int32_t local_bytes;
    int32_t max_threads;
    int32_t lds_bytes;
};

// Lane mapping: one system per lane.  Returns hipErrorInvalidValue when no
// instantiation exists for (ups, nk, field).
hipError_t lane_launch(const LaunchArgs& a, int block, hipStream_t stream);
// persistent launch of `waves` waves (queue mode)
hipError_t lane_launch_persistent(const LaunchArgs& a, int block, int waves, hipStream_t stream);
hipError_t lane_attr(int ups, int nk, int field, KernelAttr* out);
bool lane_supported(int ups, int nk, int field);

// Group mapping, one wave per workgroup: the buffers' TILE = next_pow2(ups) lanes
// per system, one per bubble (ups = 2..4, whose lanes each solve the whole coupling
// system); for ups = 5..8 two lanes per bubble (16 per system), one per coupling
// row; for ups = 9..16 one system per wave (64 lanes: the two row lanes of each
// bubble, and a copy of them that carries the bubble without taking part in the
// solve).
constexpr int group_tile(int ups) { return ups <= 1 ? 1 : 2 * group_tile((ups + 1) / 2); }
constexpr int group_lanes(int ups)
{
    return group_tile(ups) == 16 ? 64 : group_tile(ups) == 8 ? 16 : group_tile(ups);
}

hipError_t group_launch(const LaunchArgs& a, hipStream_t stream);
hipError_t group_launch_persistent(const LaunchArgs& a, int waves, hipStream_t stream);
hipError_t group_attr(int ups, int nk, int field, KernelAttr* out);
bool group_supported(int ups, int nk, int field);
int group_systems_per_block(int ups);

// Fused launch: the kernel whose bubble-count list holds every ups[b] (distinct),
// with slot[b] = the position of ups[b] in that list; null when no list covers
// the set or no kernel exists for (nk, field).
const void* fused_pick(const int* ups, int nb, int nk, int field, int* slot);
hipError_t fused_launch(const void* fn, const FusedArgs& f, int waves, hipStream_t stream);
hipError_t fused_attr(const void* fn, KernelAttr* out);

}  // namespace cbc
