// cbc_serve.h -- device side of the serving mode: the host-mapped work ring.
//
// The host appends ServeItems to a ring in coherent pinned host memory and then
// raises `ring_tail`; persistent waves claim items with a compare-and-swap on
// the device counter `queue_head`, bounded by the tail they last read, so an
// item is claimed exactly once and only after it was published.  A finished
// item's final state goes to the resident device buffers (for the next item of
// the same system) and to a host-mapped copy; a system-scope release fence then
// orders those stores before `done_seq[sys] = seq`, which the host polls.  A
// claiming wave issues a system-scope acquire fence before it loads the state,
// so it sees what another wave (on another XCD) stored for that system.
//
// Every wait is bounded: an idle wave exits when the host raises `ring_stop`
// or after `idle_ticks` wall-clock ticks without work; the host relaunches the
// kernel when items are published while none runs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cbc_kernels.h"

namespace cbc {

__device__ __forceinline__ uint32_t sys_load_u32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ int32_t sys_load_i32(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ double sys_load_f64(const double* p)
{
    const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_SYSTEM);
    return __longlong_as_double((long long)v);
}

// Load from a device buffer, or with a system-scope load from a host-mapped
// staging mirror (rewritten by the host between uses).
__device__ __forceinline__ double ld_src(const double* p, bool stage)
{
    return stage ? sys_load_f64(p) : *p;
}

__device__ __forceinline__ void sys_store_u32(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Claims up to `want` items for the calling lane (the wave's leader): reads the
// published tail once and advances queue_head by compare-and-swap; returns the
// number claimed and their first index in *base.  A failed swap means another
// wave claimed, so the loop makes progress; it is capped all the same.
__device__ __forceinline__ uint32_t serve_claim(const LaunchArgs& a, uint32_t want, uint32_t* base)
{
    const uint32_t tail = sys_load_u32(a.ring_tail);
    uint32_t head = __hip_atomic_load(a.queue_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int tries = 0; tries < 64; ++tries) {
        const uint32_t avail = tail - head;  // modulo 2^32
        if (avail == 0u || avail > 0x80000000u) return 0u;
        const uint32_t n = avail < want ? avail : want;
        if (__hip_atomic_compare_exchange_strong(a.queue_head, &head, head + n, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            *base = head;
            return n;
        }
    }
    return 0u;
}

// Item `idx` of the ring as three 16-byte loads (the 48-byte items start at
// 16-byte multiples of the page-aligned ring).  The slot is rewritten by the host
// between uses; the loads come after the claim's system-scope acquire fence,
// which orders them after the tail read that covered the item (the host writes
// the item, then publishes the tail with a release store) and invalidates
// cached lines, so they read what the host published.  Seven system-scope loads,
// one per field, each crossed the host link on their own.
__device__ __forceinline__ ServeItem serve_item(const LaunchArgs& a, uint32_t idx)
{
    static_assert(sizeof(ServeItem) == 48 && alignof(ServeItem) <= 16, "ServeItem layout");
    const uint4* p = reinterpret_cast<const uint4*>(a.ring + (idx & a.ring_mask));
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2];
    ServeItem it;
    it.sys = (int32_t)w0.x;
    it.seq = w0.y;
    it.opts = (int32_t)w0.z;
    it.pad = 0;
    it.t_end = __hiloint2double((int)w1.y, (int)w1.x);
    it.h0 = __hiloint2double((int)w1.w, (int)w1.z);
    it.pa[0] = __hiloint2double((int)w2.y, (int)w2.x);
    it.pa[1] = __hiloint2double((int)w2.w, (int)w2.z);
    return it;
}

__device__ __forceinline__ void fence_acquire_system() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, ""); }
__device__ __forceinline__ void fence_release_system() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, ""); }

// Completion of a serving item whose system stopped at load (no state loaded).
__device__ __forceinline__ void serve_publish_flag(const LaunchArgs& a, int s, uint32_t seq, int flag)
{
    double* m = a.out_misc + (size_t)s * 4;
    m[0] = 0.0;
    m[1] = 0.0;
    m[2] = (double)flag;
    m[3] = 0.0;
    fence_release_system();
    sys_store_u32(a.done_seq + s, seq);
}

// A wave that still has live systems polls the ring every
// poll_interval << shift attempts: a poll that claims no item doubles the
// interval, up to kPollBackoff doublings, and one that claims an item resets it.
// Under an offered load most such polls find the ring drained (items arrive in
// bursts, far apart on the attempt scale) and each is a read of host-mapped
// memory; a burst is still taken up within a few polls, since hundreds of waves
// poll at random phases and each claims as many items as it has free lanes.
constexpr int kPollBackoff = 3;

__device__ __forceinline__ int poll_shift_after(int shift, uint32_t got)
{
    return got > 0u ? 0 : (shift < kPollBackoff ? shift + 1 : kPollBackoff);
}

// Idle handling of a wave with no live system.  Returns true when the wave
// should exit.  `idle_since` is 0 while the wave is busy.  Every idle round
// reads the ring (tail and claim counter) before it gets here, and those reads
// of host-mapped memory are the kernels' main traffic, so the sleep between
// rounds grows to kMaxNaps x 127 x 64 clocks (~0.2 ms): with hundreds of idle
// waves polling at random phases a new item is still seen within
// microseconds.  The stop flag is read every 8th round (the host raises it
// only at destroy).
constexpr int kMaxNaps = 64;

__device__ __forceinline__ bool serve_idle(const LaunchArgs& a, uint64_t& idle_since, int& idle_rounds)
{
    const uint64_t now = wall_clock64();
    if (idle_since == 0) {
        idle_since = now | 1ull;
        idle_rounds = 0;
    } else if (now - idle_since > a.idle_ticks) {
        return true;
    }
    if ((idle_rounds & 7) == 0 && sys_load_u32(a.ring_stop) != 0u) return true;
    const int naps = idle_rounds < kMaxNaps ? idle_rounds + 1 : kMaxNaps;
    for (int k = 0; k < naps; ++k) __builtin_amdgcn_s_sleep(127);
    ++idle_rounds;
    return false;
}

}  // namespace cbc
