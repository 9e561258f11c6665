// cbc_group_kernel.h -- the group-mapped stepper kernel (one wave per workgroup,
// 64/TILE systems per wave, one lane per bubble).  Included by the per-N
// translation units cbc_group_n*.hip.
//
// Step logic is that of the reference stepper (rkck45_coupled_numba_stepper.py:179-706)
// as in cbc_lane.hip; every per-system quantity (t, h, event value, error ratio,
// flags) is held redundantly by all lanes of the group.  The whole wave keeps
// looping until every system in it has stopped, so barriers stay uniform;
// lanes of stopped systems compute but never store.
#pragma once

#include "cbc_group.h"

namespace cbc {

template <int N, int NK, int FIELD>
__global__ void __launch_bounds__(64) group_kernel(const LaunchArgs a)
{
    using Sh = GroupShape<N>;
    constexpr int G = Sh::G;
    constexpr int S = Sh::S;
    __shared__ GroupLds<N> lds[S];

    const int lane = threadIdx.x;
    const int ls = lane / G;
    const int i = lane % G;
    const int s = blockIdx.x * S + ls;
    const bool vsys = s < a.ns;
    const bool vu = vsys && i < N;
    GroupLds<N>& L = lds[ls];
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const size_t NST = NS * TILE;
    const size_t sc = vsys ? (size_t)s : 0;  // clamped index for safe loads
    const size_t base = sc * TILE;
    const int iu = vu ? i : 0;

    // load and early termination (rkck45_coupled_numba_stepper.py:179-213)
    double t = a.t[sc];
    const double t_end = a.t_end[sc];
    double h = a.h[sc];
    double g_old = a.event[sc];
    int flag = kActive;
    bool active = vsys;
    if (vsys) {
        if (t >= t_end) {
            flag = kSuccess;
            active = false;
        }
        if (fabs(g_old) < a.etol) {
            flag = kEventTerminal;
            active = false;
        }
        if (!active && i == 0) a.flags[s] = flag;
    }
    if (!__any(active)) return;  // wave-uniform

    GroupParams<NK> P;
#pragma unroll
    for (int j = 0; j < 12; ++j) P.up.u[j] = a.up[j * NST + base + iu];
    P.rcol = a.up[12 * NST + base + iu];
    P.sp0 = a.sp[sc];
    P.sp3 = a.sp[3 * NS + sc];
    {
        const double TWO_PI = 6.28318530717958647692;
        const double w0 = TWO_PI * a.sp[NS + sc];
        const double l0 = TWO_PI * a.sp[2 * NS + sc];
        const double sp4 = a.sp[4 * NS + sc];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const double amps = a.cp[(size_t)k * NS + sc];
            const double omega = a.cp[(size_t)(NK + k) * NS + sc];
            const double phase = a.cp[(size_t)(2 * NK + k) * NS + sc];
            const double wave = a.cp[(size_t)(3 * NK + k) * NS + sc];
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
    if (vu) {
        const size_t cstride = (size_t)a.ups * NST;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const size_t off = ((size_t)j * NS + sc) * TILE + i;
            L.mx[0][i][j] = a.cm[off];
            L.mx[1][i][j] = a.cm[cstride + off];
            L.mx[2][i][j] = a.cm[2 * cstride + off];
        }
    }

    double y[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) y[c] = vu ? a.state[c * NST + base + i] : (c == 0 ? 1.0 : (c == 1 ? 1e6 * (i + 1) : 0.0));
    __syncthreads();

    {
        const double g0 = group_event<N>(L, i, vu, P.rcol, y, a.col_factor);
        if (active) g_old = g0;
    }

    // dense output (rkck45_coupled_numba_stepper.py:72-112, 333-345)
    const bool dense = a.ndo > 0;
    int didx = 0;
    double dnext = t, ddt = 0.0;
    if (dense && vsys) {
        const double tb = a.t_begin[sc];
        didx = a.dense_index[sc];
        ddt = (a.ndo == 1) ? (t_end - tb) : (t_end - tb) / (a.ndo - 1);
        if (active && didx < a.ndo && dnext <= t) {
            if (i == 0) a.dense_time[(size_t)didx * NS + s] = t;
            if (vu)
#pragma unroll
                for (int c = 0; c < 4; ++c) a.dense_state[(((size_t)didx * 4 + c) * NS + s) * TILE + i] = y[c];
            ++didx;
            dnext = fmin(t + ddt, t_end);
        }
    }

    // period windows (resumable; see cbc_lane.hip)
    const bool win = a.period > 0.0;
    int wcur = 0;
    double wend = t_end;
    double wmax = y[0], wsum = 0.0;
    if (win && vsys) {
        wcur = (int)floor(t / a.period + 1e-12);
        wend = fmin((wcur + 1) * a.period, t_end);
        const int widx = wcur - a.first_window;
        if (vu && (t - wcur * a.period) > 1e-9 * a.period && widx >= 0 && widx < a.n_windows) {
            wmax = a.max_r[((size_t)widx * NS + s) * TILE + i];
            wsum = a.mean_x[((size_t)widx * NS + s) * TILE + i] * a.period;
        }
    }

    int attempts = 0, rejected = 0;
    for (int64_t step = 0; step < a.kernel_steps; ++step) {
        if (!__any(active)) break;
        if (active) ++attempts;
        if (h < a.min_step) h = a.min_step;
        if (h > a.max_step) h = a.max_step;
        {
            const double remain = wend - t;
            if (h > remain) h = remain;
        }

        double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], tmp[4];
        group_rhs<N, NK, FIELD>(P, L, i, vu, t, y, k1);
#pragma unroll
        for (int c = 0; c < 4; ++c) tmp[c] = y[c] + h * (CK::b21 * k1[c]);
        group_rhs<N, NK, FIELD>(P, L, i, vu, t + CK::a2 * h, tmp, k2);
#pragma unroll
        for (int c = 0; c < 4; ++c) tmp[c] = y[c] + h * (CK::b31 * k1[c] + CK::b32 * k2[c]);
        group_rhs<N, NK, FIELD>(P, L, i, vu, t + CK::a3 * h, tmp, k3);
#pragma unroll
        for (int c = 0; c < 4; ++c) tmp[c] = y[c] + h * (CK::b41 * k1[c] + CK::b42 * k2[c] + CK::b43 * k3[c]);
        group_rhs<N, NK, FIELD>(P, L, i, vu, t + CK::a4 * h, tmp, k4);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            tmp[c] = y[c] + h * (CK::b51 * k1[c] + CK::b52 * k2[c] + CK::b53 * k3[c] + CK::b54 * k4[c]);
        group_rhs<N, NK, FIELD>(P, L, i, vu, t + CK::a5 * h, tmp, k5);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            tmp[c] = y[c] + h * (CK::b61 * k1[c] + CK::b62 * k2[c] + CK::b63 * k3[c] + CK::b64 * k4[c] +
                                 CK::b65 * k5[c]);
        group_rhs<N, NK, FIELD>(P, L, i, vu, t + CK::a6 * h, tmp, k6);

        double ynew[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            ynew[c] = y[c] + h * (CK::c1 * k1[c] + CK::c2 * k2[c] + CK::c3 * k3[c] + CK::c4 * k4[c] +
                                  CK::c5 * k5[c] + CK::c6 * k6[c]);
        const double g_new = group_event<N>(L, i, vu, P.rcol, ynew, a.col_factor);

        // error ratio of the lane's bubble, reduced over the group
        double ratio_u = 1.0e30;
        bool bad_u = false;
        if (vu) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double e = fabs(h * (CK::e1 * k1[c] + CK::e2 * k2[c] + CK::e3 * k3[c] + CK::e4 * k4[c] +
                                           CK::e5 * k5[c] + CK::e6 * k6[c])) + 1.0e-18;
                const double tol = fmax(a.rtol * fmax(fabs(ynew[c]), fabs(y[c])), a.atol);
                const double r = tol * rcp_fast(e);
                bad_u |= (r != r) || !(fabs(ynew[c]) <= 1.7976931348623157e308);
                if (r < ratio_u) ratio_u = r;
            }
        }
        const double ratio = group_min<G>(ratio_u);
        const bool bad = group_any<G>(bad_u);

        if (active) {
            const bool crossing = g_old * g_new < 0.0;
            bool stop = false;
            if (crossing) {
                if (h <= a.min_step) {
                    flag = kStepSizeLimit;
                    stop = true;
                } else {
                    const double aold = fabs(g_old);
                    h = h * ((aold * rcp_fast(aold + fabs(g_new) + 1.0e-15)) * 0.99);
                    ++rejected;
                }
            } else if (bad) {
                flag = kNaN;
                stop = true;
            } else {
                const bool ok = ratio >= 1.0;
                if (ok ? (t + h == t) : (h <= a.min_step)) {
                    flag = kStepSizeLimit;
                    stop = true;
                } else {
                    const double fac = step_factor(ok, ratio);
                    if (ok) {
                        if (win) {
                            if (ynew[0] > wmax) wmax = ynew[0];
                            wsum += 0.5 * h * (y[1] + ynew[1]);
                        }
                        t = t + h;
#pragma unroll
                        for (int c = 0; c < 4; ++c) y[c] = ynew[c];
                        g_old = g_new;
                    } else {
                        ++rejected;
                    }
                    h = h * fac;
                }
            }
            if (stop) {
                active = false;
            } else {
                if (dense && didx < a.ndo && dnext <= t) {
                    if (i == 0) a.dense_time[(size_t)didx * NS + s] = t;
                    if (vu)
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            a.dense_state[(((size_t)didx * 4 + c) * NS + s) * TILE + i] = y[c];
                    ++didx;
                    dnext = fmin(t + ddt, t_end);
                }
                if (win && t >= wend) {
                    const int widx = wcur - a.first_window;
                    if (vu && widx >= 0 && widx < a.n_windows) {
                        a.max_r[((size_t)widx * NS + s) * TILE + i] = wmax;
                        a.mean_x[((size_t)widx * NS + s) * TILE + i] = wsum / a.period;
                    }
                    ++wcur;
                    wend = fmin((wcur + 1) * a.period, t_end);
                    wmax = y[0];
                    wsum = 0.0;
                }
                if (t >= t_end) {
                    flag = kSuccess;
                    active = false;
                } else if (fabs(g_old) < a.etol) {
                    flag = kEventTerminal;
                    active = false;
                }
            }
        }
    }

    if (!vsys) return;
    // systems that were stopped at load have nothing to store
    if (attempts == 0 && flag != kActive) return;
    if (win && flag != kSuccess) {
        const int widx = wcur - a.first_window;
        if (vu && (t - wcur * a.period) > 1e-9 * a.period && widx >= 0 && widx < a.n_windows) {
            a.max_r[((size_t)widx * NS + s) * TILE + i] = wmax;
            a.mean_x[((size_t)widx * NS + s) * TILE + i] = wsum / a.period;
        }
    }
    if (vu)
#pragma unroll
        for (int c = 0; c < 4; ++c) a.state[c * NST + base + i] = y[c];
    if (i == 0) {
        a.t[s] = t;
        a.h[s] = h;
        a.event[s] = g_old;
        a.flags[s] = flag;
        if (dense) a.dense_index[s] = didx;
        if (a.total_steps) a.total_steps[s] += attempts;
        if (a.rejected_steps) a.rejected_steps[s] += rejected;
        if (flag == kActive && a.active_count) atomicAdd(a.active_count, 1u);
    }
}

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
