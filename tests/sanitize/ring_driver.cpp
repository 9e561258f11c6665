// ring_driver.cpp -- exercises the serving ring's host logic (cbc_ring.h) against a
// simulated device, for the sanitizer test (tests/test_sanitizers_cpu.py builds it
// with -fsanitize=address,undefined).  The simulated device claims items in ring
// order and finishes them in random order, writing done_seq as the serving kernel
// does.  Checks: a slot is retired only after its item and every older one
// finished; every submission check (ring full, index range, unfinished item,
// duplicates, option bits) rejects without writing; sequence numbers wrap past
// 2^32; ring indices wrap past the capacity.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../coupled-bubble-control_amd/csrc/cbc_ring.h"

using namespace cbc;

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                                     \
        }                                                                     \
    } while (0)

int main()
{
    const int ns = 100;
    const uint32_t cap = 64;
    std::vector<ServeItem> ring(cap);
    std::vector<uint32_t> done(ns), last(ns);
    // start every system's sequence numbers just below 2^32 so they wrap
    for (int s = 0; s < ns; ++s) done[s] = last[s] = 0xFFFFFFF0u + (uint32_t)(s % 7);
    RingHost r{ring.data(), cap, ns, 2, done.data(), last.data(), 0xFFFFFFC0u, 0xFFFFFFC0u};
    std::mt19937 rng(7);
    std::vector<uint32_t> claimed;      // ring positions claimed by the device, not yet finished
    uint32_t head = r.tail;             // device claim counter
    long submitted = 0, finished = 0, rejected_full = 0;
    for (int round = 0; round < 20000; ++round) {
        // host: submit a few idle systems (some submissions deliberately invalid)
        const int kind = (int)(rng() % 16);
        std::vector<int32_t> sys;
        for (int k = 0, want = 1 + (int)(rng() % 6); k < want; ++k) {
            const int32_t q = (int32_t)(rng() % ns);
            bool seen = false;
            for (int32_t x : sys) seen = seen || x == q;
            if (!seen && done[q] == last[q]) sys.push_back(q);
        }
        if (sys.empty()) sys.push_back((int32_t)(rng() % ns));
        std::vector<int32_t> opts(sys.size(), (int32_t)(rng() % 4 == 0 ? kServeReset : 0));
        if (kind == 0) sys.push_back(sys[0]);                       // duplicate
        if (kind == 1) sys.push_back(ns + (int32_t)(rng() % 5));    // out of range
        if (kind == 2) opts[0] |= 8;                                // unknown option bit
        const int n = (int)sys.size();
        opts.resize(n, 0);
        std::vector<double> pa(2 * n, 1.0e5), h0(n, 0.0);
        std::vector<uint32_t> seq(n, 0);
        const std::vector<uint32_t> last_before = last;
        const uint32_t tail_before = r.tail;
        int32_t bad = 0;
        const RingStatus st = ring_submit(r, sys.data(), pa.data(), n, 5.0, h0.data(), opts.data(), seq.data(), &bad);
        if (st != kRingOk) {
            CHECK(r.tail == tail_before);
            CHECK(last == last_before);
            if (st == kRingFull) ++rejected_full;
            if (st == kRingDuplicate) CHECK(kind == 0);
            if (st == kRingBadSystem) CHECK(kind == 1 && (bad < 0 || bad >= ns));
            if (st == kRingBadOptions) CHECK(kind == 2);
            if (st == kRingBusySystem) CHECK(done[bad] != last[bad]);
        } else {
            CHECK(kind > 2);                                        // every invalid submission is rejected
            CHECK(r.tail - r.retired <= cap);
            for (int k = 0; k < n; ++k) {
                const ServeItem& it = ring[(tail_before + (uint32_t)k) & (cap - 1)];
                CHECK(it.sys == sys[k] && it.seq == seq[k] && it.seq == last[sys[k]]);
                CHECK(it.seq == last_before[sys[k]] + 1u);
            }
            submitted += n;
        }
        // device: claim everything published, finish a random claimed item
        while (head != r.tail) claimed.push_back(head++);
        for (int k = 0, m = (int)(rng() % 4); k < m && !claimed.empty(); ++k) {
            const size_t j = rng() % claimed.size();
            const ServeItem& it = ring[claimed[j] & (cap - 1)];
            CHECK((int32_t)(it.seq - done[it.sys]) == 1);       // items of a system finish in order
            __atomic_store_n(&done[it.sys], it.seq, __ATOMIC_RELEASE);
            claimed.erase(claimed.begin() + (long)j);
            ++finished;
        }
        // host: retire; every retired slot's item must be finished, the next one not
        ring_retire(r);
        for (uint32_t p : claimed) CHECK((int32_t)(p - r.retired) >= 0);
        if (r.retired != r.tail) CHECK(!ring_item_done(r, ring[r.retired & (cap - 1)]));
    }
    std::printf("ok: %ld items submitted, %ld finished, %ld full-ring rejections, tail %u\n", submitted, finished,
                rejected_full, r.tail);
    return 0;
}
