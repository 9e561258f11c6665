// cbc_ring.h -- host side of the serving work ring: the item layout, slot
// retirement and submission checks.  Plain C++ (no HIP), so the logic is also
// built with host sanitizers on its own (tests/sanitize/ring_driver.cpp).
//
// The ring lives in coherent pinned host memory that the serving kernel reads;
// done_seq is host-mapped memory the kernel writes (per system, the sequence
// number of its last finished item).  Slots are freed in submission order: slot
// `retired` is reused only after its item and every older one have finished.
#pragma once

#include <stdint.h>

namespace cbc {

// ServeItem.opts bits
enum : int32_t {
    kServeReset = 1,        // new episode: state, unit parameters and coupling rows come from the staging mirrors
    kServeUseControl = 2,   // amplitudes from control_params (set_action's rows) instead of the item
};

// One env step of one system: integrate from the resident state (actual_state,
// time_step) over tau in [0, t_end] with drive amplitudes pa[0..NK) [Pa]; the
// stored event value is not consulted (the environment resets it every step).
// h0 > 0 replaces the resident step size.  With kServeReset the system first
// takes its state, unit parameters and coupling rows from the staging mirrors
// and stores them into the device buffers.  48 bytes, written by the host only.
struct ServeItem {
    int32_t sys;
    uint32_t seq;
    int32_t opts;
    int32_t pad;
    double t_end;
    double h0;
    double pa[2];
};
static_assert(sizeof(ServeItem) == 48, "ServeItem layout");

struct RingHost {
    ServeItem* ring;          // capacity slots (a power of two)
    uint32_t capacity;
    int32_t ns;               // systems
    int32_t nk;               // drive amplitudes per item (1 or 2)
    const uint32_t* done_seq; // (ns) written by the device
    uint32_t* last_seq;       // (ns) host: seq of the last item submitted per system
    uint32_t tail;            // items written
    uint32_t retired;         // oldest slot not yet known finished
};

inline uint32_t ring_load_acquire(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }

// An item is finished when done_seq[sys] has reached its seq (modulo 2^32).
inline bool ring_item_done(const RingHost& r, const ServeItem& it)
{
    return (int32_t)(ring_load_acquire(&r.done_seq[it.sys]) - it.seq) >= 0;
}

// Frees the slots of finished items, oldest first, stopping at the first item
// still running.
inline void ring_retire(RingHost& r)
{
    while (r.retired != r.tail) {
        const ServeItem& it = r.ring[r.retired & (r.capacity - 1)];
        if (!ring_item_done(r, it)) break;
        ++r.retired;
    }
}

enum RingStatus : int32_t {
    kRingOk = 0,
    kRingFull = 1,          // fewer free slots than items: nothing written
    kRingBadSystem = 2,     // index out of range (err_sys)
    kRingBusySystem = 3,    // the system has an unfinished item (err_sys)
    kRingDuplicate = 4,     // the system appears twice in one submission (err_sys)
    kRingBadOptions = 5,    // unknown option bits (err_sys = the bits)
};

// Writes n items into the ring after checking them; on success advances `tail`
// (the caller then publishes it to the device) and fills seq_out.  On failure
// nothing is written and last_seq is unchanged.
inline RingStatus ring_submit(RingHost& r, const int32_t* sys, const double* pa, int32_t n, double t_end,
                              const double* h0, const int32_t* opts, uint32_t* seq_out, int32_t* err_sys)
{
    ring_retire(r);
    const uint32_t used = r.tail - r.retired;
    if ((int64_t)used + n > (int64_t)r.capacity) return kRingFull;
    for (int32_t k = 0; k < n; ++k) {
        const int32_t q = sys[k];
        if (q < 0 || q >= r.ns) {
            *err_sys = q;
            return kRingBadSystem;
        }
        if (ring_load_acquire(&r.done_seq[q]) != r.last_seq[q]) {
            *err_sys = q;
            return kRingBusySystem;
        }
        if (opts && (opts[k] & ~(kServeReset | kServeUseControl)) != 0) {
            *err_sys = opts[k];
            return kRingBadOptions;
        }
    }
    // duplicates within this submission: advance last_seq as items are taken and
    // undo on the first system seen twice
    for (int32_t k = 0; k < n; ++k) {
        const int32_t q = sys[k];
        if (r.last_seq[q] != ring_load_acquire(&r.done_seq[q])) {
            for (int32_t j = 0; j < k; ++j) --r.last_seq[sys[j]];
            *err_sys = q;
            return kRingDuplicate;
        }
        ++r.last_seq[q];
    }
    for (int32_t k = 0; k < n; ++k) {
        const int32_t q = sys[k];
        ServeItem& it = r.ring[(r.tail + (uint32_t)k) & (r.capacity - 1)];
        it.sys = q;
        it.seq = r.last_seq[q];
        it.opts = opts ? opts[k] : 0;
        it.pad = 0;
        it.t_end = t_end;
        it.h0 = h0 ? h0[k] : 0.0;
        it.pa[0] = pa[k];
        it.pa[1] = r.nk > 1 ? pa[(size_t)n + k] : 0.0;
        if (seq_out) seq_out[k] = it.seq;
    }
    r.tail += (uint32_t)n;
    return kRingOk;
}

}  // namespace cbc
