// cbc_lane.hip -- lane mapping: one CKM1D system per lane of a wave64.
//
// The whole system (4N state values, the six Cash--Karp stages, the unit and
// coupling constants and the 2N x 2N coupling matrix) lives in VGPRs; there
// is no LDS, no barrier and no cross-lane traffic, so the lanes of a wave
// step independent systems with independent step sizes.  The per-system
// algorithm is that of the reference stepper (rkck45_coupled_numba_stepper.py:179-706,
// rkck45_coupled_cuda_sover.cu:274-845), which runs one *thread per bubble*
// and synchronises the block ~20 times per step; here a system never leaves
// its lane.
#include "cbc_device.h"
#include "cbc_linalg.h"
#include "cbc_kernels.h"
#include "cbc_serve.h"

namespace cbc {

template <int N, int NK>
struct LaneSys {
    Unit up[N];
    double rcol[N];           // up[12] = R0 / lambda_ref, collision radius scale
    double m0[N][N], m1[N][N], m2[N][N];  // m_c[i][j]: effect of bubble j on bubble i
    double sp0, sp3;          // 3*kappa, liquid sound speed
    Drive<NK> d;
};

// Harmonics and polytropic factors of one stage of one system.
template <int N, int NK>
struct LaneStage {
    double st[NK], ct[NK];        // time harmonics (shared by the bubbles)
    double sx[N][NK], cx[N][NK];  // space harmonics per bubble
    double pf[N];                 // (R/R0)^(-3 kappa) per bubble
};

// First stage of an attempt, in full; rxb = 1/(R/R0).
template <int N, int NK, int FIELD>
__device__ __forceinline__ void lane_stage_base(const LaneSys<N, NK>& P, double t, const double (&y)[4 * N],
                                                LaneStage<N, NK>& B, double (&rxb)[N])
{
    time_base<NK>(P.d, t, B.st, B.ct);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if constexpr (FIELD != kFieldConst) space_base<NK>(P.d, y[N + i], B.sx[i], B.cx[i]);
        B.pf[i] = pow_pos(y[i], -P.sp0);
        rxb[i] = rcp_fast(y[i]);
    }
}

// A later stage at (tq, x) from the first stage B at (t, y).
template <int N, int NK, int FIELD>
__device__ __forceinline__ void lane_stage_next(const LaneSys<N, NK>& P, const LaneStage<N, NK>& B,
                                                const double (&rxb)[N], double t, const double (&y)[4 * N],
                                                double tq, const double (&x)[4 * N], LaneStage<N, NK>& S)
{
    time_stage<NK>(P.d, B.st, B.ct, t, tq, true, S.st, S.ct);
    if constexpr (FIELD != kFieldConst) {
        bool far = false;
#pragma unroll
        for (int i = 0; i < N; ++i) far |= space_far<NK>(P.d, y[N + i], x[N + i]);
        if (far) {
#pragma unroll
            for (int i = 0; i < N; ++i) space_base<NK>(P.d, x[N + i], S.sx[i], S.cx[i]);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) space_shift<NK>(P.d, B.sx[i], B.cx[i], y[N + i], x[N + i], S.sx[i], S.cx[i]);
        }
    }
    bool far = false;
#pragma unroll
    for (int i = 0; i < N; ++i) far |= pf_far(P.sp0, y[i], rxb[i], x[i]);
    if (far) {
#pragma unroll
        for (int i = 0; i < N; ++i) S.pf[i] = pow_pos(x[i], -P.sp0);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) S.pf[i] = pf_shift(P.sp0, B.pf[i], y[i], rxb[i], x[i]);
    }
}

// Full coupled RHS of one system from its stage values; y, dy ordered
// [R(N), x(N), U(N), V(N)].
template <int N, int NK, int FIELD>
__device__ __forceinline__ void lane_rhs(const LaneSys<N, NK>& P, const LaneStage<N, NK>& S,
                                         const double (&y)[4 * N], double (&dy)[4 * N])
{
    BubbleOut o[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double pa, pat, pax, uax;
        field_from<NK, FIELD>(P.d, S.st, S.ct, S.sx[i], S.cx[i], pa, pat, pax, uax);
        bubble_rhs(P.up[i], P.sp3, y[i], y[2 * N + i], y[3 * N + i], pa, pat, pax, uax, S.pf[i], o[i]);
        dy[i] = y[2 * N + i];
        dy[N + i] = y[3 * N + i];
    }

    if constexpr (N == 1) {
        dy[2] = o[0].dU;
        dy[3] = o[0].dV;
    } else {
        // pair geometry, one division per unordered pair
        double rd[N][N], sg[N][N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
#pragma unroll
            for (int j = i + 1; j < N; ++j) {
                const double delta = y[N + i] - y[N + j];
                const double s = sgn(delta);
                const double r = rcp_fast(fmax(fabs(delta), 1e-30));
                rd[i][j] = r; rd[j][i] = r;
                sg[i][j] = s; sg[j][i] = -s;
            }
        }
        // explicit coupling (ckm1d_def.cuh:103-166) and the implicit system
        // (ckm1d_def.cuh:213-248) assembled as [I + C | b]
        double a[2 * N][2 * N + 1];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double rg0 = 0.0, rg1 = 0.0, tng = 0.0, tg2 = 0.0, tg3 = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (j == i) {
                    a[i][j] = 1.0;
                    a[i][N + j] = 0.0;
                    a[N + i][j] = 0.0;
                    a[N + i][N + j] = 1.0;
                    continue;
                }
                const double r = rd[i][j];
                const double r2 = r * r;
                const double sr2 = r2 * sg[i][j];
                const double r3 = r2 * r;
                const double m0 = P.m0[i][j], m1 = P.m1[i][j], m2 = P.m2[i][j];
                rg0 += m0 * (-2.0 * r * o[j].H0 - 2.5 * sr2 * o[j].H2 - 1.0 * r3 * o[j].H4);
                rg1 += m0 * (-0.5 * sr2 * o[j].H1 - 0.5 * r3 * o[j].H3);
                tng += m1 * (2.0 * sr2 * o[j].H0 + 5.0 * r3 * o[j].H2);
                tg2 += m1 * (sr2 * o[j].H1 + r3 * o[j].H3);
                tg3 += m2 * (sr2 * o[j].H1 + r3 * o[j].H3);
                a[i][j] = o[i].G0 * m0 * r * o[j].H5;
                a[i][N + j] = 0.5 * o[i].G0 * m0 * sr2 * o[j].H6;
                a[N + i][j] = -m1 * sr2 * o[j].H5;
                a[N + i][N + j] = -m1 * r3 * o[j].H6;
            }
            a[i][2 * N] = o[i].dU + (rg0 * o[i].G0 + rg1 * o[i].G1);
            a[N + i][2 * N] = o[i].dV + (tng + tg2 * o[i].G2 + tg3 * o[i].G3);
        }
        double v[2 * N];
        lu_solve<2 * N>(a, v);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dy[2 * N + i] = v[i];
            dy[3 * N + i] = v[N + i];
        }
    }
}

// Collision gap min_i |x_i - x_{i+1}| - colf (r_i + r_{i+1}), r = R/R0 * R0/lambda;
// ckm1d_def.py:512-552.  A single bubble has no gap (1e30).
template <int N, int NK>
__device__ __forceinline__ double lane_event(const LaneSys<N, NK>& P, const double (&y)[4 * N], double colf)
{
    double g = 1e30;
#pragma unroll
    for (int i = 0; i + 1 < N; ++i) {
        const double ri = y[i] * P.rcol[i];
        const double rj = y[i + 1] * P.rcol[i + 1];
        const double gap = fabs(y[N + i] - y[N + i + 1]) - (ri + rj) * colf;
        if (gap < g) g = gap;
    }
    return g;
}

// Root theta of the gap along the step y0 -> y1 of length h (exact event mode).
template <int N, int NK>
__device__ double lane_event_root(const LaneSys<N, NK>& P, const double (&y0)[4 * N], const double (&y1)[4 * N],
                                  double h, double colf, double g0, double g1)
{
    Illinois st{0.0, g0, 1.0, g1, false};
    for (int it = 0; it < kIllinoisIters && !st.done; ++it) {
        const double c = illinois_point(st);
        double g = 1e30;
        double r[N], x[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            r[i] = hermite3(y0[i], y1[i], h * y0[2 * N + i], h * y1[2 * N + i], c) * P.rcol[i];
            x[i] = hermite3(y0[N + i], y1[N + i], h * y0[3 * N + i], h * y1[3 * N + i], c);
        }
#pragma unroll
        for (int i = 0; i + 1 < N; ++i) {
            const double gap = fabs(x[i] - x[i + 1]) - (r[i] + r[i + 1]) * colf;
            if (gap < g) g = gap;
        }
        illinois_update(st, c, g);
    }
    return st.b;
}

// Per-lane integration state of one system.
template <int N, int NK>
struct LaneRun {
    LaneSys<N, NK> P;
    double y[4 * N];
    double t, t_end, h, g_old;
    int s, flag, attempts, rejected;
    uint32_t seq;             // serving: sequence number of the current item
    int didx;
    double dnext, ddt;
    int wcur;
    double wend;
    double wmax[N], wsum[N];
};

// Load system s (rkck45_coupled_numba_stepper.py:179-213, 248-358).  Returns false
// when the system is already finished at load (its flag is written then).  A
// serving item (`it` != null) starts the system at tau = 0 with the item's end
// time, amplitudes and optional initial step, ignoring the stored event value
// (the environment's per-step reset, envs/bubble_1D_pos_control.py:731-749, 815-826).
template <int N, int NK, int FIELD>
__device__ __forceinline__ bool lane_begin(const LaunchArgs& a, int s, LaneRun<N, NK>& R,
                                           const ServeItem* it = nullptr)
{
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const size_t NST = NS * TILE;
    const size_t base = (size_t)s * TILE;
    R.s = s;
    // a failure stop is final within one solve call (the host loop of the
    // oracle stops such a system for good): later launches leave it untouched
    if (!it && a.resume) {
        const int f = a.flags[s];
        if (f < 0) {
            R.flag = f;
            return false;
        }
    }
    R.t = it ? 0.0 : a.t[s];
    R.t_end = it ? it->t_end : a.t_end[s];
    R.h = (it && it->h0 > 0.0) ? it->h0 : a.h[s];
    R.g_old = it ? -1.0 : a.event[s];
    R.flag = kActive;
    R.attempts = 0;
    R.rejected = 0;
    {
        bool term = R.t >= R.t_end;
        int flag = term ? kSuccess : kActive;
        if (fabs(R.g_old) < a.etol) {
            term = true;
            flag = kEventTerminal;
        }
        if (term) {
            a.flags[s] = flag;
            R.flag = flag;
            return false;
        }
    }
    LaneSys<N, NK>& P = R.P;
    // a new episode (serving reset) reads its rows from the host-mapped staging
    // mirrors and stores them into the device buffers
    const bool rs = it && (it->opts & kServeReset);
    const double* src_up = rs ? a.stage_up : a.up;
    const double* src_cm = rs ? a.stage_cm : a.cm;
    const double* src_y = rs ? a.stage_state : a.state;
    const size_t cstride = (size_t)a.ups * NST;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < 12; ++j) P.up[i].u[j] = ld_src(src_up + j * NST + base + i, rs);
        P.rcol[i] = ld_src(src_up + 12 * NST + base + i, rs);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            // coupling[c][j][s][i] is the effect of unit j on unit i (ckm1d_def.cuh:56-57)
            const size_t off = ((size_t)j * NS + s) * TILE + i;
            P.m0[i][j] = ld_src(src_cm + off, rs);
            P.m1[i][j] = ld_src(src_cm + cstride + off, rs);
            P.m2[i][j] = ld_src(src_cm + 2 * cstride + off, rs);
        }
    }
    if (rs) {
        double* up = const_cast<double*>(a.up);
        double* cm = const_cast<double*>(a.cm);
#pragma unroll
        for (int i = 0; i < N; ++i) {
#pragma unroll
            for (int j = 0; j < 12; ++j) up[j * NST + base + i] = P.up[i].u[j];
            up[12 * NST + base + i] = P.rcol[i];
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const size_t off = ((size_t)j * NS + s) * TILE + i;
                cm[off] = P.m0[i][j];
                cm[cstride + off] = P.m1[i][j];
                cm[2 * cstride + off] = P.m2[i][j];
            }
        }
    }
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
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) R.y[j * N + i] = ld_src(src_y + j * NST + base + i, rs);

    R.g_old = lane_event<N, NK>(P, R.y, a.col_factor);

    // dense output (rkck45_coupled_numba_stepper.py:72-112, 333-345)
    R.didx = 0;
    R.dnext = R.t;
    R.ddt = 0.0;
    if (a.ndo > 0) {
        const double tb = a.t_begin[s];
        R.didx = a.dense_index[s];
        R.ddt = (a.ndo == 1) ? (R.t_end - tb) : (R.t_end - tb) / (a.ndo - 1);
        if (R.didx < a.ndo && R.dnext <= R.t) {
            a.dense_time[(size_t)R.didx * NS + s] = R.t;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < N; ++i)
                    a.dense_state[(((size_t)R.didx * 4 + j) * NS + s) * TILE + i] = R.y[j * N + i];
            ++R.didx;
            R.dnext = fmin(R.t + R.ddt, R.t_end);
        }
    }

    // period windows: a window end is an intermediate end time of the clamp.
    // A launch that starts inside a window resumes its partial statistics,
    // which the previous launch left in max_r / mean_x.
    R.wcur = 0;
    R.wend = R.t_end;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        R.wmax[i] = R.y[i];
        R.wsum[i] = 0.0;
    }
    if (a.period > 0.0) {
        R.wcur = (int)floor(R.t / a.period + 1e-12);
        R.wend = fmin((R.wcur + 1) * a.period, R.t_end);
        const int widx = R.wcur - a.first_window;
        const bool mid = (R.t - R.wcur * a.period) > 1e-9 * a.period;
        if (mid && widx >= 0 && widx < a.n_windows) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                R.wmax[i] = a.max_r[((size_t)widx * NS + s) * TILE + i];
                R.wsum[i] = a.mean_x[((size_t)widx * NS + s) * TILE + i] * a.period;
            }
        }
    }
    return true;
}

// One Cash--Karp attempt (rkck45_coupled_numba_stepper.py:362-677).  Returns true
// when the system stopped (R.flag set).
template <int N, int NK, int FIELD>
__device__ __forceinline__ bool lane_attempt(const LaunchArgs& a, LaneRun<N, NK>& R)
{
    const LaneSys<N, NK>& P = R.P;
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const int s = R.s;
    double& t = R.t;
    double& h = R.h;
    double (&y)[4 * N] = R.y;
    ++R.attempts;
    // clamp (rkck45_coupled_numba_stepper.py:366-373)
    if (h < a.min_step) h = a.min_step;
    if (h > a.max_step) h = a.max_step;
    {
        const double remain = R.wend - t;
        if (h > remain) h = remain;
    }

    // Cash--Karp stages (rkck45_coupled_numba_stepper.py:379-484): k_q goes at
    // once into the inputs of the later stages (acc[r]), the 5th-order sum and
    // the error sum, in the reference's order of summation.  The loop is
    // unrolled: a rolled loop (one copy of the RHS, smaller code) measured
    // slower on gfx950 because the accumulators then stay live across the
    // loop and spill to AGPRs.
    double acc[6][4 * N], ysum[4 * N], esum[4 * N], x[4 * N], k[4 * N];
#pragma unroll
    for (int c = 0; c < 4 * N; ++c) {
#pragma unroll
        for (int r = 1; r < 6; ++r) acc[r][c] = 0.0;
        ysum[c] = 0.0;
        esum[c] = 0.0;
    }
    LaneStage<N, NK> B, S;
    double rxb[N];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (q == 0) {
#pragma unroll
            for (int c = 0; c < 4 * N; ++c) x[c] = y[c];
            lane_stage_base<N, NK, FIELD>(P, t, y, B, rxb);
            S = B;
        } else {
#pragma unroll
            for (int r = 1; r < 6; ++r) {
                if (r == q) {
#pragma unroll
                    for (int c = 0; c < 4 * N; ++c) x[c] = y[c] + h * acc[r][c];
                }
            }
            lane_stage_next<N, NK, FIELD>(P, B, rxb, t, y, t + kCkA[q] * h, x, S);
        }
        lane_rhs<N, NK, FIELD>(P, S, x, k);
#pragma unroll
        for (int r = 1; r < 6; ++r) {
            if (r > q) {
                const double b = kCkB[r][q];
#pragma unroll
                for (int c = 0; c < 4 * N; ++c) acc[r][c] += b * k[c];
            }
        }
        const double cq = kCkC[q], eq = kCkE[q];
#pragma unroll
        for (int c = 0; c < 4 * N; ++c) {
            ysum[c] += cq * k[c];
            esum[c] += eq * k[c];
        }
    }

    double ynew[4 * N];
#pragma unroll
    for (int c = 0; c < 4 * N; ++c) ynew[c] = y[c] + h * ysum[c];
    const double g_new = lane_event<N, NK>(P, ynew, a.col_factor);

    // event crossing (rkck45_coupled_numba_stepper.py:500-515, 575-581); in the
    // exact mode a crossing step that lands within etol of the root is accepted
    const bool crossing = R.g_old * g_new < 0.0 && !(a.event_exact && fabs(g_new) < a.etol);
    if (crossing) {
        // cannot shrink further: no progress possible; a NaN step size (the retry
        // factor of an infinite gap) passes every clamp and would never end
        if (!(h > a.min_step)) {
            R.flag = h == h ? kStepSizeLimit : kNaN;
            return true;
        }
        double ratio;
        if (a.event_exact) {
            ratio = fmin(lane_event_root<N, NK>(P, y, ynew, h, a.col_factor, R.g_old, g_new), kEventThetaMax);
        } else {
            const double aold = fabs(R.g_old);
            ratio = (aold * rcp_fast(aold + fabs(g_new) + 1.0e-15)) * 0.99;
        }
        h = h * ratio;
        ++R.rejected;
    } else {
        // error norm (rkck45_coupled_numba_stepper.py:525-536)
        double ratio = 1.0e30;
        bool bad = false;
#pragma unroll
        for (int c = 0; c < 4 * N; ++c) {
            const double e = fabs(h * esum[c]) + 1.0e-18;
            const double tol = fmax(a.rtol * fmax(fabs(ynew[c]), fabs(y[c])), a.atol);
            const double r = tol * rcp_fast(e);
            bad |= (r != r) || !(fabs(ynew[c]) <= 1.7976931348623157e308);
            if (r < ratio) ratio = r;
        }
        if (bad) {
            R.flag = kNaN;
            return true;
        }
        const bool ok = ratio >= 1.0;
        // step size below the time resolution: an accepted step would not
        // advance t, a rejected one cannot shrink (the reference clamps h at
        // MIN_STEP and spins until its step budget runs out)
        if (ok ? (t + h == t) : (h <= a.min_step)) {
            R.flag = kStepSizeLimit;
            return true;
        }
        const double fac = step_factor(ok, ratio);
        if (ok) {
            if (a.period > 0.0) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    if (ynew[i] > R.wmax[i]) R.wmax[i] = ynew[i];
                    R.wsum[i] += 0.5 * h * (y[N + i] + ynew[N + i]);
                }
            }
            t = t + h;
#pragma unroll
            for (int c = 0; c < 4 * N; ++c) y[c] = ynew[c];
            R.g_old = g_new;
        } else {
            ++R.rejected;
        }
        h = h * fac;
    }

    if (a.ndo > 0 && R.didx < a.ndo && R.dnext <= t) {
        a.dense_time[(size_t)R.didx * NS + s] = t;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < N; ++i)
                a.dense_state[(((size_t)R.didx * 4 + j) * NS + s) * TILE + i] = y[j * N + i];
        ++R.didx;
        R.dnext = fmin(t + R.ddt, R.t_end);
    }

    if (a.period > 0.0 && t >= R.wend) {
        const int widx = R.wcur - a.first_window;
        if (widx >= 0 && widx < a.n_windows) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                a.max_r[((size_t)widx * NS + s) * TILE + i] = R.wmax[i];
                a.mean_x[((size_t)widx * NS + s) * TILE + i] = R.wsum[i] / a.period;
            }
        }
        ++R.wcur;
        R.wend = fmin((R.wcur + 1) * a.period, R.t_end);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            R.wmax[i] = y[i];
            R.wsum[i] = 0.0;
        }
    }

    // termination (rkck45_coupled_numba_stepper.py:651-663)
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

// Store (rkck45_coupled_numba_stepper.py:683-704); also saves an open window.
template <int N, int NK, int FIELD>
__device__ __forceinline__ void lane_finish(const LaunchArgs& a, const LaneRun<N, NK>& R)
{
    const size_t NS = (size_t)a.ns;
    const size_t TILE = (size_t)a.tile;
    const size_t NST = NS * TILE;
    const int s = R.s;
    const size_t base = (size_t)s * TILE;
    if (a.period > 0.0 && R.flag != kSuccess) {
        const int widx = R.wcur - a.first_window;
        if ((R.t - R.wcur * a.period) > 1e-9 * a.period && widx >= 0 && widx < a.n_windows) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                a.max_r[((size_t)widx * NS + s) * TILE + i] = R.wmax[i];
                a.mean_x[((size_t)widx * NS + s) * TILE + i] = R.wsum[i] / a.period;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) a.state[j * NST + base + i] = R.y[j * N + i];
    a.t[s] = R.t;
    a.h[s] = R.h;
    a.event[s] = R.g_old;
    a.flags[s] = R.flag;
    if (a.ndo > 0) a.dense_index[s] = R.didx;
    if (a.total_steps) a.total_steps[s] += R.attempts;
    if (a.rejected_steps) a.rejected_steps[s] += R.rejected;
    if (R.flag == kActive && a.active_count) atomicAdd(a.active_count, 1u);
    if (a.ring) {
        if (a.out_state)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < N; ++i) a.out_state[j * NST + base + i] = R.y[j * N + i];
        double* m = a.out_misc + (size_t)s * 4;
        m[0] = R.t;
        m[1] = R.h;
        m[2] = (double)R.flag;
        m[3] = (double)R.attempts;
        fence_release_system();
        sys_store_u32(a.done_seq + s, R.seq);
    }
}

// Plain mode (queue == null): lane = system, at most kernel_steps attempts.
// Persistent mode: each wave refills its free lanes from the work queue with one
// atomic per refill (free lanes take consecutive queue slots in lane order) once
// at least `refill` lanes are free, or all are; the wave exits when the queue is
// drained and its lanes are idle.  Serving mode: as persistent mode, but items
// come from the host-mapped ring (polled every `poll_interval` attempts while
// some lane is busy, up to 8x as often after polls that found nothing:
// poll_shift_after), and an idle wave waits for work until the host stops it
// or `idle_ticks` pass.  All modes run the same attempt loop.
template <int N, int NK, int FIELD>
__global__ void __launch_bounds__(256) lane_kernel(const LaunchArgs a)
{
    LaneRun<N, NK> R;
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
        const int s = blockIdx.x * blockDim.x + threadIdx.x;
        if (s < a.ns) busy = lane_begin<N, NK, FIELD>(a, s, R);
    }
    const int lane = threadIdx.x & 63;
    while (true) {
        if (!drained) {
            const uint64_t freemask = __ballot(!busy);
            const int nfree = __popcll(freemask);
            bool take = nfree > 0 && (nfree >= a.refill || nfree == 64);
            if (serving) take = take && (nfree == 64 || since_poll >= (a.poll_interval << poll_shift));
            if (take) {
                const int leader = __ffsll((unsigned long long)freemask) - 1;
                const uint32_t rank = (uint32_t)__popcll(freemask & ((1ull << lane) - 1ull));
                if (serving) {
                    since_poll = 0;
                    uint32_t base = 0, got = 0;
                    if (lane == leader) got = serve_claim(a, (uint32_t)nfree, &base);
                    got = __shfl(got, leader);
                    base = __shfl(base, leader);
                    poll_shift = poll_shift_after(poll_shift, got);
                    if (got > 0) fence_acquire_system();
                    if (!busy && rank < got) {
                        const ServeItem it = serve_item(a, base + rank);
                        if (it.sys >= 0 && it.sys < a.ns) {  // the host validates; never trust the ring
                            busy = lane_begin<N, NK, FIELD>(a, it.sys, R, &it);
                            R.seq = it.seq;
                            used = 0;
                            if (!busy) serve_publish_flag(a, it.sys, it.seq, R.flag);
                        }
                    }
                } else {
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(a.queue_head, (uint32_t)nfree);
                    base = __shfl(base, leader);
                    if (base + (uint32_t)nfree >= (uint32_t)a.queue_len) drained = true;
                    if (!busy) {
                        const uint32_t idx = base + rank;
                        if (idx < (uint32_t)a.queue_len) {
                            const int q = a.queue[idx];
                            if (q >= 0 && q < a.ns) {  // malformed entries are skipped
                                busy = lane_begin<N, NK, FIELD>(a, q, R);
                                used = 0;
                            }
                        }
                    }
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
        if (busy) {
            const bool stop = lane_attempt<N, NK, FIELD>(a, R);
            if (stop || ++used >= a.kernel_steps) {
                lane_finish<N, NK, FIELD>(a, R);
                busy = false;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <int N, int NK, int FIELD>
static const void* lane_fn()
{
    return reinterpret_cast<const void*>(&lane_kernel<N, NK, FIELD>);
}

template <int N, int NK>
static const void* lane_pick_field(int field)
{
    switch (field) {
    case kFieldConst: return lane_fn<N, NK, kFieldConst>();
    case kFieldSwN: return lane_fn<N, NK, kFieldSwN>();
    case kFieldSwA: return lane_fn<N, NK, kFieldSwA>();
    default: return nullptr;
    }
}

template <int N>
static const void* lane_pick_nk(int nk, int field)
{
    switch (nk) {
    case 1: return lane_pick_field<N, 1>(field);
    case 2: return lane_pick_field<N, 2>(field);
    default: return nullptr;
    }
}

static const void* lane_pick(int ups, int nk, int field)
{
    switch (ups) {
    case 1: return lane_pick_nk<1>(nk, field);
    case 2: return lane_pick_nk<2>(nk, field);
    case 3: return lane_pick_nk<3>(nk, field);
    case 4: return lane_pick_nk<4>(nk, field);
    default: return nullptr;
    }
}

bool lane_supported(int ups, int nk, int field) { return lane_pick(ups, nk, field) != nullptr; }

hipError_t lane_launch(const LaunchArgs& a, int block, hipStream_t stream)
{
    const void* fn = lane_pick(a.ups, a.nk, a.field);
    if (!fn) return hipErrorInvalidValue;
    if (block <= 0 || block > 256 || (block % 64) != 0) return hipErrorInvalidValue;
    const int grid = (a.ns + block - 1) / block;
    void* args[] = {const_cast<LaunchArgs*>(&a)};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, stream);
}

hipError_t lane_launch_persistent(const LaunchArgs& a, int block, int waves, hipStream_t stream)
{
    const void* fn = lane_pick(a.ups, a.nk, a.field);
    if (!fn || !(a.queue || a.ring) || !a.queue_head) return hipErrorInvalidValue;
    if (block <= 0 || block > 256 || (block % 64) != 0 || waves <= 0) return hipErrorInvalidValue;
    const int wpb = block / 64;
    const int grid = (waves + wpb - 1) / wpb;
    void* args[] = {const_cast<LaunchArgs*>(&a)};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, stream);
}

hipError_t lane_attr(int ups, int nk, int field, KernelAttr* out)
{
    const void* fn = lane_pick(ups, nk, field);
    if (!fn) return hipErrorInvalidValue;
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, fn);
    if (e != hipSuccess) return e;
    out->num_regs = fa.numRegs;
    out->local_bytes = (int32_t)fa.localSizeBytes;
    out->max_threads = fa.maxThreadsPerBlock;
    out->lds_bytes = (int32_t)fa.sharedSizeBytes;
    return hipSuccess;
}

}  // namespace cbc
