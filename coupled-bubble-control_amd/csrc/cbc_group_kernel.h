// cbc_group_kernel.h -- the group-mapped stepper kernel (one wave per workgroup,
// 64/TILE systems per wave, one lane per bubble).  Included by the per-N
// translation units cbc_group_n*.hip.
//
// Step logic is that of the reference stepper (rkck45_coupled_numba_stepper.py:179-706)
// as in cbc_lane.hip; every per-system quantity (t, h, event value, error ratio,
// flags) is held redundantly by all lanes of the group.  The whole wave runs
// every attempt together (the RHS syncs and reduces across the wave), so lanes of idle groups
// compute but never store.
//
// Plain mode: group = system blockIdx.x * S + group index.  Lane i of a group
// carries bubble i; in the groups of N = 5..16, lanes i and TILE + i both carry
// bubble i (its radial and its translational coupling row; for N = 9..16 lanes
// 32 + i and 48 + i a copy of them) and step it identically, and lane i alone
// stores its results.
// Persistent mode: free groups take systems from the work queue (one atomic per
// refill by the wave, consecutive slots to free groups in lane order) once at
// least `refill` groups are free, or all are.
#pragma once

#include "cbc_group.h"
#include "cbc_serve.h"

namespace cbc {

template <int N, int NK>
struct GroupRun {
    GroupParams<NK> P;
    double y[4];
    double t, t_end, h, g_old;
    int s, flag, attempts, rejected;
    uint32_t seq;  // serving: sequence number of the current item
    bool vu;  // lane carries a real bubble of a loaded system
    int didx;
    double dnext, ddt;
    int wcur;
    double wend, wmax, wsum;
};

// Per-lane load of system s into group slot L (no barrier).  Returns whether the
// system is active; a system finished at load gets its flag written.  A serving
// item (`it` != null) starts at tau = 0 as in lane_begin (cbc_lane.hip).
template <int N, int NK>
__device__ __forceinline__ bool group_begin(const LaunchArgs& a, int s, int i, bool wr, GroupLds<N>& L,
                                            GroupRun<N, NK>& R, const ServeItem* it = nullptr)
{
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const size_t NST = NS * TILE;
    const size_t base = (size_t)s * TILE;
    R.s = s;
    // a failure stop is final within one solve call: later launches leave it untouched
    if (!it && a.resume) {
        const int f = a.flags[s];
        if (f < 0) {
            R.flag = f;
            R.vu = false;
            return false;
        }
    }
    R.vu = i < N;
    const int iu = R.vu ? i : 0;
    R.t = it ? 0.0 : a.t[s];
    R.t_end = it ? it->t_end : a.t_end[s];
    R.h = (it && it->h0 > 0.0) ? it->h0 : a.h[s];
    R.g_old = it ? -1.0 : a.event[s];
    R.flag = kActive;
    R.attempts = 0;
    R.rejected = 0;
    bool active = true;
    if (R.t >= R.t_end) {
        R.flag = kSuccess;
        active = false;
    }
    if (fabs(R.g_old) < a.etol) {
        R.flag = kEventTerminal;
        active = false;
    }
    if (!active) {
        if (i == 0 && wr) a.flags[s] = R.flag;
        R.vu = false;
        return false;
    }
    GroupParams<NK>& P = R.P;
    // a new episode (serving reset) reads its rows from the host-mapped staging
    // mirrors and stores them into the device buffers
    const bool rs = it && (it->opts & kServeReset);
    const double* src_up = rs ? a.stage_up : a.up;
    const double* src_cm = rs ? a.stage_cm : a.cm;
    const double* src_y = rs ? a.stage_state : a.state;
#pragma unroll
    for (int j = 0; j < 12; ++j) P.up.u[j] = ld_src(src_up + j * NST + base + iu, rs);
    P.rcol = ld_src(src_up + 12 * NST + base + iu, rs);
    P.sp0 = a.sp[s];
    P.sp3 = a.sp[3 * NS + s];
    {
        const double TWO_PI = 6.28318530717958647692;
        const double w0 = TWO_PI * a.sp[NS + s];
        const double l0 = TWO_PI * a.sp[2 * NS + s];
        const double sp4 = a.sp[4 * NS + s];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const double amps = (it && !(it->opts & kServeUseControl)) ? it->pa[k] : a.cp[(size_t)k * NS + s];
            const double omega = a.cp[(size_t)(NK + k) * NS + s];
            const double phase = a.cp[(size_t)(2 * NK + k) * NS + s];
            const double wave = a.cp[(size_t)(3 * NK + k) * NS + s];
            P.d.wt[k] = w0 * omega;
            P.d.wx[k] = l0 * wave;
            P.d.ph[k] = phase;
            P.d.A[k] = amps;
            P.d.Aw[k] = amps * omega;
            P.d.Ak[k] = amps * wave;
            P.d.Au[k] = amps * sp4;
        }
    }
    // coupling constants: coupling[c][j][s][i] is the effect of unit j on unit i
    if (R.vu && wr) {
        const size_t cstride = (size_t)a.ups * NST;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const size_t off = ((size_t)j * NS + s) * TILE + i;
            L.mx[0][i][j] = ld_src(src_cm + off, rs);
            L.mx[1][i][j] = ld_src(src_cm + cstride + off, rs);
            L.mx[2][i][j] = ld_src(src_cm + 2 * cstride + off, rs);
        }
        if (rs) {
            double* up = const_cast<double*>(a.up);
            double* cm = const_cast<double*>(a.cm);
#pragma unroll
            for (int j = 0; j < 12; ++j) up[j * NST + base + i] = P.up.u[j];
            up[12 * NST + base + i] = P.rcol;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const size_t off = ((size_t)j * NS + s) * TILE + i;
                cm[off] = L.mx[0][i][j];
                cm[cstride + off] = L.mx[1][i][j];
                cm[2 * cstride + off] = L.mx[2][i][j];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
        R.y[c] = R.vu ? ld_src(src_y + c * NST + base + i, rs) : (c == 0 ? 1.0 : (c == 1 ? 1e6 * (i + 1) : 0.0));

    // dense output (rkck45_coupled_numba_stepper.py:72-112, 333-345)
    R.didx = 0;
    R.dnext = R.t;
    R.ddt = 0.0;
    if (a.ndo > 0) {
        const double tb = a.t_begin[s];
        R.didx = a.dense_index[s];
        R.ddt = (a.ndo == 1) ? (R.t_end - tb) : (R.t_end - tb) / (a.ndo - 1);
        if (R.didx < a.ndo && R.dnext <= R.t) {
            if (i == 0 && wr) a.dense_time[(size_t)R.didx * NS + s] = R.t;
            if (R.vu && wr)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    a.dense_state[(((size_t)R.didx * 4 + c) * NS + s) * TILE + i] = R.y[c];
            ++R.didx;
            R.dnext = fmin(R.t + R.ddt, R.t_end);
        }
    }
    // period windows (resumable; see cbc_lane.hip)
    R.wcur = 0;
    R.wend = R.t_end;
    R.wmax = R.y[0];
    R.wsum = 0.0;
    if (a.period > 0.0) {
        R.wcur = (int)floor(R.t / a.period + 1e-12);
        R.wend = fmin((R.wcur + 1) * a.period, R.t_end);
        const int widx = R.wcur - a.first_window;
        if (R.vu && (R.t - R.wcur * a.period) > 1e-9 * a.period && widx >= 0 && widx < a.n_windows) {
            R.wmax = a.max_r[((size_t)widx * NS + s) * TILE + i];
            R.wsum = a.mean_x[((size_t)widx * NS + s) * TILE + i] * a.period;
        }
    }
    return true;
}

// One attempt for every group of the wave (uniform: wave-wide syncs).  `busy`
// marks groups with a live system; returns true for a busy group whose system
// stopped this attempt.
template <int N, int NK, int FIELD>
__device__ __forceinline__ bool group_attempt(const LaunchArgs& a, int i, int half, GroupLds<N>& L,
                                              GroupRun<N, NK>& R, bool busy)
{
    constexpr int T = GroupShape<N>::T;  // per-bubble reductions: over one half of a split group
    const bool wr = half == 0;
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    double& t = R.t;
    double& h = R.h;
    double (&y)[4] = R.y;
    const bool vu = busy && R.vu;
    if (busy) ++R.attempts;
    if (h < a.min_step) h = a.min_step;
    if (h > a.max_step) h = a.max_step;
    {
        const double remain = R.wend - t;
        if (h > remain) h = remain;
    }

    // stage loop as in lane_attempt (cbc_lane.hip)
    double acc[6][4], ysum[4], esum[4], x[4], k[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int r = 1; r < 6; ++r) acc[r][c] = 0.0;
        ysum[c] = 0.0;
        esum[c] = 0.0;
    }
    GroupStage<NK> B, S;
    double rxb;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (q == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) x[c] = y[c];
            group_stage_base<NK, FIELD>(R.P, t, y, B, rxb);
            S = B;
        } else {
#pragma unroll
            for (int r = 1; r < 6; ++r) {
                if (r == q) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) x[c] = y[c] + h * acc[r][c];
                }
            }
            group_stage_next<NK, FIELD>(R.P, B, rxb, t, y, t + kCkA[q] * h, x, vu, S);
        }
        group_rhs<N, NK, FIELD>(R.P, L, i, half, vu, S, x, k);
#pragma unroll
        for (int r = 1; r < 6; ++r) {
            if (r > q) {
                const double b = kCkB[r][q];
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] += b * k[c];
            }
        }
        const double cq = kCkC[q], eq = kCkE[q];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            ysum[c] += cq * k[c];
            esum[c] += eq * k[c];
        }
    }

    double ynew[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ynew[c] = y[c] + h * ysum[c];
    const double g_new = group_event<N>(L, i, wr, vu, R.P.rcol, ynew, a.col_factor);

    // error ratio of the lane's bubble, reduced over the group
    double ratio_u = 1.0e30;
    bool bad_u = false;
    if (vu) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double e = fabs(h * esum[c]) + 1.0e-18;
            const double tol = fmax(a.rtol * fmax(fabs(ynew[c]), fabs(y[c])), a.atol);
            const double r = tol * rcp_fast(e);
            bad_u |= (r != r) || !(fabs(ynew[c]) <= 1.7976931348623157e308);
            if (r < ratio_u) ratio_u = r;
        }
    }
    const double ratio = group_min<T>(ratio_u);
    const bool bad = group_any<T>(bad_u);
    // exact event mode: the gap's root on the step, for groups that cross
    double theta = 1.0;
    if (a.event_exact) {
        const bool need = busy && R.g_old * g_new < 0.0 && !(fabs(g_new) < a.etol);
        if (__any(need))
            theta = group_event_root<N>(L, i, wr, vu, R.P.rcol, y, ynew, h, a.col_factor, R.g_old, g_new);
    }
    if (!busy) return false;

    const int s = R.s;
    const bool crossing = R.g_old * g_new < 0.0 && !(a.event_exact && fabs(g_new) < a.etol);
    if (crossing) {
        // no progress possible; a NaN step size (the retry factor of an infinite gap)
        // passes every clamp and would be retried until the attempt budget runs out
        if (!(h > a.min_step)) {
            R.flag = h == h ? kStepSizeLimit : kNaN;
            return true;
        }
        if (a.event_exact) {
            h = h * fmin(theta, kEventThetaMax);
        } else {
            const double aold = fabs(R.g_old);
            h = h * ((aold * rcp_fast(aold + fabs(g_new) + 1.0e-15)) * 0.99);
        }
        ++R.rejected;
    } else {
        if (bad) {
            R.flag = kNaN;
            return true;
        }
        const bool ok = ratio >= 1.0;
        if (ok ? (t + h == t) : (h <= a.min_step)) {
            R.flag = kStepSizeLimit;
            return true;
        }
        const double fac = step_factor(ok, ratio);
        if (ok) {
            if (a.period > 0.0) {
                if (ynew[0] > R.wmax) R.wmax = ynew[0];
                R.wsum += 0.5 * h * (y[1] + ynew[1]);
            }
            t = t + h;
#pragma unroll
            for (int c = 0; c < 4; ++c) y[c] = ynew[c];
            R.g_old = g_new;
        } else {
            ++R.rejected;
        }
        h = h * fac;
    }
    if (a.ndo > 0 && R.didx < a.ndo && R.dnext <= t) {
        if (i == 0 && wr) a.dense_time[(size_t)R.didx * NS + s] = t;
        if (R.vu && wr)
#pragma unroll
            for (int c = 0; c < 4; ++c) a.dense_state[(((size_t)R.didx * 4 + c) * NS + s) * TILE + i] = y[c];
        ++R.didx;
        R.dnext = fmin(t + R.ddt, R.t_end);
    }
    if (a.period > 0.0 && t >= R.wend) {
        const int widx = R.wcur - a.first_window;
        if (R.vu && wr && widx >= 0 && widx < a.n_windows) {
            a.max_r[((size_t)widx * NS + s) * TILE + i] = R.wmax;
            a.mean_x[((size_t)widx * NS + s) * TILE + i] = R.wsum / a.period;
        }
        ++R.wcur;
        R.wend = fmin((R.wcur + 1) * a.period, R.t_end);
        R.wmax = y[0];
        R.wsum = 0.0;
    }
    if (t >= R.t_end) {
        R.flag = kSuccess;
        return true;
    }
    if (fabs(R.g_old) < a.etol) {
        R.flag = kEventTerminal;
        return true;
    }
    return false;
}

template <int N, int NK>
__device__ __forceinline__ void group_finish(const LaunchArgs& a, int i, bool wr, const GroupRun<N, NK>& R)
{
    if (!wr) return;  // the second lane of a bubble (split groups) stores nothing
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const size_t NST = NS * TILE;
    const int s = R.s;
    const size_t base = (size_t)s * TILE;
    if (a.period > 0.0 && R.flag != kSuccess) {
        const int widx = R.wcur - a.first_window;
        if (R.vu && (R.t - R.wcur * a.period) > 1e-9 * a.period && widx >= 0 && widx < a.n_windows) {
            a.max_r[((size_t)widx * NS + s) * TILE + i] = R.wmax;
            a.mean_x[((size_t)widx * NS + s) * TILE + i] = R.wsum / a.period;
        }
    }
    if (R.vu)
#pragma unroll
        for (int c = 0; c < 4; ++c) a.state[c * NST + base + i] = R.y[c];
    if (i == 0) {
        a.t[s] = R.t;
        a.h[s] = R.h;
        a.event[s] = R.g_old;
        a.flags[s] = R.flag;
        if (a.ndo > 0) a.dense_index[s] = R.didx;
        if (a.total_steps) a.total_steps[s] += R.attempts;
        if (a.rejected_steps) a.rejected_steps[s] += R.rejected;
        if (R.flag == kActive && a.active_count) atomicAdd(a.active_count, 1u);
    }
    if (a.ring) {
        if (R.vu && a.out_state)
#pragma unroll
            for (int c = 0; c < 4; ++c) a.out_state[c * NST + base + i] = R.y[c];
        if (i == 0) {
            double* m = a.out_misc + (size_t)s * 4;
            m[0] = R.t;
            m[1] = R.h;
            m[2] = (double)R.flag;
            m[3] = (double)R.attempts;
        }
        fence_release_system();
        if (i == 0) sys_store_u32(a.done_seq + s, R.seq);
    }
}

// The stepper of one wave (one workgroup): plain mode steps the systems of
// block `block`; persistent mode (queue or ring) takes systems until the queue
// is drained (or the serving ring idles out).  `lds` holds the wave's S group slots.
template <int N, int NK, int FIELD>
__device__ __forceinline__ void group_run(const LaunchArgs& a, int block, GroupLds<N>* lds)
{
    using Sh = GroupShape<N>;
    constexpr int G = Sh::G;
    constexpr int S = Sh::S;

    const int lane = threadIdx.x;
    const int ls = lane / G;
    const int gi = lane % G;      // lane within the group
    const int i = gi % Sh::T;     // its bubble
    const int half = gi / Sh::T;  // split groups: 0 radial-row lane, 1 translational-row lane
    const bool wr = half == 0;    // the lane that stores the bubble's results
    GroupLds<N>& L = lds[ls];
    GroupRun<N, NK> R;
    R.vu = false;
    R.y[0] = 1.0;
    R.y[1] = 1e6 * (i + 1);
    R.y[2] = 0.0;
    R.y[3] = 0.0;
    R.t = 0.0;
    R.h = 1e-3;
    R.wend = 1.0;
    R.P.rcol = 0.0;
    R.P.sp0 = 4.2;
    R.P.sp3 = 1.0;
#pragma unroll
    for (int j = 0; j < 12; ++j) R.P.up.u[j] = 1.0;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        R.P.d.wt[k] = R.P.d.wx[k] = R.P.d.ph[k] = 0.0;
        R.P.d.A[k] = R.P.d.Aw[k] = R.P.d.Ak[k] = R.P.d.Au[k] = 0.0;
    }

    bool busy = false;
    int64_t used = 0;
    const bool serving = a.ring != nullptr;
    const bool persistent = a.queue != nullptr || serving;
    bool drained = !persistent;
    int since_poll = 0x40000000;
    int poll_shift = 0;
    uint64_t idle_since = 0;
    int idle_rounds = 0;
    if (!persistent) {
        const int s = block * S + ls;
        if (s < a.ns) busy = group_begin<N, NK>(a, s, i, wr, L, R);
        __syncthreads();
        const double g0 = group_event<N>(L, i, wr, busy && R.vu, R.P.rcol, R.y, a.col_factor);
        if (busy) R.g_old = g0;
    }
    while (true) {
        if (!drained) {
            const bool free_leader = !busy && gi == 0;
            const uint64_t freemask = __ballot(free_leader);
            const int nfree = __popcll(freemask);
            const bool any_busy = __any(busy);
            bool take = nfree > 0 && (nfree >= a.refill || !any_busy);
            if (serving) take = take && (!any_busy || since_poll >= (a.poll_interval << poll_shift));
            if (take) {
                const int leader = __ffsll((unsigned long long)freemask) - 1;
                const uint32_t rank = (uint32_t)__popcll(freemask & ((1ull << lane) - 1ull));
                bool fresh = false;
                if (serving) {
                    since_poll = 0;
                    uint32_t base = 0, got = 0;
                    if (lane == leader) got = serve_claim(a, (uint32_t)nfree, &base);
                    got = __shfl(got, leader);
                    base = __shfl(base, leader);
                    poll_shift = poll_shift_after(poll_shift, got);
                    if (got > 0) fence_acquire_system();
                    int slot = (free_leader && rank < got) ? (int)rank : -1;
                    slot = __shfl(slot, ls * G);  // group leader's slot to the whole group
                    if (!busy && slot >= 0) {
                        const ServeItem it = serve_item(a, base + (uint32_t)slot);
                        if (it.sys >= 0 && it.sys < a.ns) {  // the host validates; never trust the ring
                            fresh = group_begin<N, NK>(a, it.sys, i, wr, L, R, &it);
                            R.seq = it.seq;
                            used = 0;
                            if (!fresh && gi == 0) serve_publish_flag(a, it.sys, it.seq, R.flag);
                        }
                    }
                } else {
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(a.queue_head, (uint32_t)nfree);
                    base = __shfl(base, leader);
                    if (base + (uint32_t)nfree >= (uint32_t)a.queue_len) drained = true;
                    uint32_t idx = free_leader ? base + rank : 0u;
                    idx = __shfl(idx, ls * G);  // group leader's slot to the whole group
                    const int q = (!busy && idx < (uint32_t)a.queue_len) ? a.queue[idx] : -1;
                    if (q >= 0 && q < a.ns) {  // malformed entries are skipped
                        fresh = group_begin<N, NK>(a, q, i, wr, L, R);
                        used = 0;
                    }
                }
                __syncthreads();
                const double g0 = group_event<N>(L, i, wr, fresh && R.vu, R.P.rcol, R.y, a.col_factor);
                if (fresh) {
                    R.g_old = g0;
                    busy = true;
                }
            }
        }
        if (!__any(busy)) {
            if (serving) {
                if (serve_idle(a, idle_since, idle_rounds)) break;
                continue;
            }
            if (drained) break;
            continue;
        }
        idle_since = 0;
        ++since_poll;
        const bool stop = group_attempt<N, NK, FIELD>(a, i, half, L, R, busy);
        if (busy && (stop || ++used >= a.kernel_steps)) {
            group_finish<N, NK>(a, i, wr, R);
            busy = false;
            R.vu = false;
        }
    }
}

template <int N, int NK, int FIELD>
__global__ void __launch_bounds__(64) group_kernel(const LaunchArgs a)
{
    __shared__ GroupLds<N> lds[GroupShape<N>::S];
    group_run<N, NK, FIELD>(a, blockIdx.x, lds);
}

// Every build carries the exact event block and tests the mode at run time.
template <int N>
const void* group_pick(int nk, int field)
{
#define CBC_G(NK_, F_) return reinterpret_cast<const void*>(&group_kernel<N, NK_, F_>)
    if (nk == 1) {
        if (field == kFieldConst) CBC_G(1, kFieldConst);
        if (field == kFieldSwN) CBC_G(1, kFieldSwN);
        if (field == kFieldSwA) CBC_G(1, kFieldSwA);
    } else if (nk == 2) {
        if (field == kFieldConst) CBC_G(2, kFieldConst);
        if (field == kFieldSwN) CBC_G(2, kFieldSwN);
        if (field == kFieldSwA) CBC_G(2, kFieldSwA);
    }
#undef CBC_G
    return nullptr;
}

}  // namespace cbc
