// cbc_device.h -- gfx950 device code shared by the CKM1D kernels.
//
// The coupled Keller--Miksis 1-D model (CKM1D) of the reference:
//   per-bubble RHS      backends/cupy_cuda/cuda_src/ckm1d_def.cuh:379-490,
//                       backends/numba/sysmtem_definitions/ckm1d_def.py:407-496
//   acoustic field      ckm1d_def.py:40-155 (particle-velocity prefactor sp[4])
//   explicit coupling   ckm1d_def.cuh:66-178
//   implicit coupling   ckm1d_def.cuh:180-376, solved here by LU with partial
//                       pivoting instead of the reference's Richardson sweep
//   collision event     ckm1d_def.py:498-556
//   Cash--Karp tableau  rkck45_coupled_cuda_sover.cu:137-147
// All arithmetic is FP64; expression order follows the reference text so that
// the CPU restatement and the kernels agree to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cbc_math.h"

namespace cbc {

// status flags (rkck45_coupled_numba_stepper.py:27-33)
enum : int {
    kActive = 0,
    kSuccess = 1,
    kEventTerminal = 2,
    kMaxSteps = -1,
    kStepSizeLimit = -2,
    kConvergence = -3,
    kNaN = -4,
};

enum : int { kFieldConst = 0, kFieldSwN = 1, kFieldSwA = 2 };

// Cash--Karp coefficients.  A[j]: stage times; B[q][j]: weight of k_j in the
// input of stage q; C: 5th-order weights; E = C - C*, the embedded error.
struct CK {
    static constexpr double a2 = 1.0 / 5.0, a3 = 3.0 / 10.0, a4 = 3.0 / 5.0, a5 = 1.0, a6 = 7.0 / 8.0;
    static constexpr double b21 = 1.0 / 5.0;
    static constexpr double b31 = 3.0 / 40.0, b32 = 9.0 / 40.0;
    static constexpr double b41 = 3.0 / 10.0, b42 = -9.0 / 10.0, b43 = 6.0 / 5.0;
    static constexpr double b51 = -11.0 / 54.0, b52 = 5.0 / 2.0, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0;
    static constexpr double b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
                            b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0;
    static constexpr double c1 = 37.0 / 378.0, c2 = 0.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0,
                            c5 = 0.0, c6 = 512.0 / 1771.0;
    static constexpr double c1p = 2825.0 / 27648.0, c2p = 0.0, c3p = 18575.0 / 48384.0,
                            c4p = 13525.0 / 55296.0, c5p = 277.0 / 14336.0, c6p = 1.0 / 4.0;
    static constexpr double e1 = c1 - c1p, e2 = c2 - c2p, e3 = c3 - c3p, e4 = c4 - c4p,
                            e5 = c5 - c5p, e6 = c6 - c6p;
};

// The tableau by stage index q = 0..5 for the rolled stage loop: kCkA[q] stage
// time fraction, kCkB[r][q] weight of k_q in the input of stage r (0 unless
// q < r), kCkC / kCkE the 5th-order and embedded-error weights.
__device__ constexpr double kCkA[6] = {0.0, CK::a2, CK::a3, CK::a4, CK::a5, CK::a6};
__device__ constexpr double kCkB[6][6] = {
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {CK::b21, 0.0, 0.0, 0.0, 0.0, 0.0},
    {CK::b31, CK::b32, 0.0, 0.0, 0.0, 0.0},
    {CK::b41, CK::b42, CK::b43, 0.0, 0.0, 0.0},
    {CK::b51, CK::b52, CK::b53, CK::b54, 0.0, 0.0},
    {CK::b61, CK::b62, CK::b63, CK::b64, CK::b65, 0.0}};
__device__ constexpr double kCkC[6] = {CK::c1, CK::c2, CK::c3, CK::c4, CK::c5, CK::c6};
__device__ constexpr double kCkE[6] = {CK::e1, CK::e2, CK::e3, CK::e4, CK::e5, CK::e6};

__device__ __forceinline__ double sgn(double v) { return (double)((v > 0.0) - (v < 0.0)); }

// Per-system constants of the acoustic drive, folded once per launch from the
// control parameters [PA(NK), omega(NK), phase(NK), wave(NK)] (models/ckm1d.py:190-202)
// and the system parameters (ckm1d.py:171-175).
template <int NK>
struct Drive {
    double wt[NK];  // (2*pi*sp[1]) * omega   -> time argument factor
    double wx[NK];  // (2*pi*sp[2]) * wave    -> space argument factor
    double ph[NK];  // phase
    double A[NK];   // amplitude
    double Aw[NK];  // amplitude * omega
    double Ak[NK];  // amplitude * wave
    double Au[NK];  // amplitude * sp[4]
};

// Pressure, its time derivative, its gradient and the particle velocity at
// (t, x) given the time harmonics st/ct[k] = sin/cos(wt_k t + ph_k) and the
// space harmonics sx/cx[k] = sin/cos(wx_k x + ph_k); ckm1d_def.py:40-155.
template <int NK, int FIELD>
__device__ __forceinline__ void field_from(const Drive<NK>& d, const double (&st)[NK], const double (&ct)[NK],
                                           const double (&sx)[NK], const double (&cx)[NK],
                                           double& p, double& pt, double& px, double& ux)
{
    p = 0.0; pt = 0.0; px = 0.0; ux = 0.0;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        if constexpr (FIELD == kFieldConst) {
            p += d.A[k] * st[k];
            pt += d.Aw[k] * ct[k];
        } else if constexpr (FIELD == kFieldSwN) {
            p += d.A[k] * sx[k] * st[k];
            pt += d.Aw[k] * sx[k] * ct[k];
            px += d.Ak[k] * cx[k] * st[k];
            ux += d.Au[k] * cx[k] * ct[k];
        } else {
            p += d.A[k] * cx[k] * st[k];
            pt += d.Aw[k] * cx[k] * ct[k];
            px -= d.Ak[k] * sx[k] * st[k];
            ux -= d.Au[k] * sx[k] * ct[k];
        }
    }
}

// Stage values reused within one attempt.  The first stage (t, x, x0 = R/R0)
// is evaluated in full; later stages (tq, xq, x0q) start from it by the
// addition theorems and the (1+e)^y series when the lane's own arguments are
// inside their bounds, else in full.  The choice is made per lane, never per
// wave, so a system's arithmetic does not depend on the systems sharing its
// wave.  Lanes with vote = false (no live bubble) always take the series
// path, so they never make the wave run both.

// time harmonics sin/cos(wt_k t + ph_k)
template <int NK>
__device__ __forceinline__ void time_base(const Drive<NK>& d, double t, double (&st)[NK], double (&ct)[NK])
{
#pragma unroll
    for (int k = 0; k < NK; ++k) sincos_fast(d.wt[k] * t + d.ph[k], &st[k], &ct[k]);
}

template <int NK>
__device__ __forceinline__ void time_stage(const Drive<NK>& d, const double (&sb0)[NK], const double (&cb0)[NK],
                                           double t, double tq, bool vote, double (&st)[NK], double (&ct)[NK])
{
    const double dt = tq - t;
    double b[NK];
    bool far = false;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        b[k] = d.wt[k] * dt;
        far |= !(fabs(b[k]) <= kPi4);
    }
    if (vote && far) {
        time_base<NK>(d, tq, st, ct);
    } else {
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            double sb, cm1;
            sincos_kern(b[k], &sb, &cm1);
            rotate(sb0[k], cb0[k], sb, cm1, &st[k], &ct[k]);
        }
    }
}

// space harmonics sin/cos(wx_k x + ph_k) of one bubble
template <int NK>
__device__ __forceinline__ void space_base(const Drive<NK>& d, double x, double (&sx)[NK], double (&cx)[NK])
{
#pragma unroll
    for (int k = 0; k < NK; ++k) sincos_fast(d.wx[k] * x + d.ph[k], &sx[k], &cx[k]);
}

template <int NK>
__device__ __forceinline__ bool space_far(const Drive<NK>& d, double xb, double xq)
{
    const double dx = xq - xb;
    bool far = false;
#pragma unroll
    for (int k = 0; k < NK; ++k) far |= !(fabs(d.wx[k] * dx) <= kTinyAngle);
    return far;
}

template <int NK>
__device__ __forceinline__ void space_shift(const Drive<NK>& d, const double (&sb0)[NK], const double (&cb0)[NK],
                                            double xb, double xq, double (&sx)[NK], double (&cx)[NK])
{
    const double dx = xq - xb;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        double sb, cm1;
        sincos_tiny(d.wx[k] * dx, &sb, &cm1);
        rotate(sb0[k], cb0[k], sb, cm1, &sx[k], &cx[k]);
    }
}

// polytropic factor (R/R0)^(-sp0) from its first-stage value pf0 at x0b (rxb = 1/x0b)
__device__ __forceinline__ bool pf_far(double sp0, double x0b, double rxb, double x0q)
{
    return !(fabs((x0q - x0b) * rxb) <= kSmallRatio) || !(fabs(sp0) <= kSmallExp);
}

__device__ __forceinline__ double pf_shift(double sp0, double pf0, double x0b, double rxb, double x0q)
{
    return pf0 * pow1p_small((x0q - x0b) * rxb, -sp0);
}

// Per-bubble constants: up[0..11] of models/ckm1d.py:153-167.
struct Unit {
    double u[12];
};

// Uncoupled radial (Keller--Miksis) and translational RHS of one bubble plus
// its coupling factors G and coupling terms H; ckm1d_def.cuh:401-452.
struct BubbleOut {
    double dU, dV;
    double G0, G1, G2, G3;
    double H0, H1, H2, H3, H4, H5, H6;
};

__device__ __forceinline__ void bubble_rhs(const Unit& up, double sp3, double x0, double x2, double x3,
                                           double pa, double pat, double pax, double uax, double p,
                                           BubbleOut& o)
{
    // p = (1/x0)^sp0, the polytropic factor (Harm::pf)
    const double rx0 = rcp_fast(x0);
    const double N = (up.u[0] + up.u[1] * x2) * p
                   - up.u[2] * (1 + up.u[7] * x2)
                   - up.u[3] * rx0
                   - up.u[4] * x2 * rx0
                   - 1.5 * (1.0 - up.u[7] * x2 * (1.0 / 3.0)) * x2 * x2
                   - (1 + up.u[7] * x2) * up.u[5] * pa
                   - up.u[6] * pat * x0
                   + up.u[8] * x3 * x3;
    const double D = x0 - up.u[7] * x0 * x2 + up.u[4] * up.u[7];
    const double rD = rcp_fast(D);
    const double Fb1 = -up.u[10] * x0 * x0 * x0 * pax;
    const double Fd = -up.u[11] * x0 * (x3 * sp3 - uax);
    o.dU = N * rD;
    o.dV = 3 * (Fb1 + Fd) * up.u[9] * rx0 * rx0 * rx0 - 3.0 * x2 * rx0 * x3;
    o.G0 = rD;
    o.G1 = x3 * rD;
    o.G2 = x2 * rx0;
    o.G3 = rx0 * rx0;
    const double x02 = x0 * x0;
    const double x03 = x02 * x0;
    const double x22 = x2 * x2;
    const double x32 = x3 * x3;
    o.H0 = x0 * x22;
    o.H1 = x02 * x2;
    o.H2 = x02 * x2 * x3;
    o.H3 = x03 * x3;
    o.H4 = x03 * x32;
    o.H5 = x02;
    o.H6 = x03;
}

// Step-size factor of rkck45_coupled_numba_stepper.py:610-613:
//   accepted: min(5, max(0.2, 0.9 ratio^0.2)), rejected: min(0.5, max(0.1, 0.9 ratio^0.25)).
__device__ __forceinline__ double step_factor(bool ok, double ratio)
{
    const double f = 0.9 * pow_pos(ratio, ok ? 0.2 : 0.25);
    return ok ? fmin(5.0, fmax(0.2, f)) : fmin(0.5, fmax(0.1, f));
}

// ---------------------------------------------------------------------------
// Exact collision-event localisation (CBC_EVENT_EXACT).  Along one step of
// length h the radii and positions follow the cubic Hermite interpolant of their
// end values with slopes h U and h V (R' = U and x' = V in the dimensionless
// model), so the gap is known on the whole step without another RHS
// evaluation; its root theta in (0, 1) is found by the Illinois method
// (modified regula falsi, Dowell & Jarratt 1971) from the bracket
// (0, g_old), (1, g_new).  The iteration count is fixed at most and a
// converged iterate is frozen, so every mapping and the oracle reach the same
// theta from the same states.

constexpr int kIllinoisIters = 60;
constexpr double kEventThetaMax = 0.99;   // a retried step always shrinks

__device__ __forceinline__ double hermite3(double p0, double p1, double d0, double d1, double th)
{
    const double t2 = th * th, t3 = t2 * th;
    const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0;
    const double h10 = t3 - 2.0 * t2 + th;
    const double h01 = 3.0 * t2 - 2.0 * t3;
    const double h11 = t3 - t2;
    return h00 * p0 + h10 * d0 + h01 * p1 + h11 * d1;
}

struct Illinois {
    double a, fa, b, fb;
    bool done;
};

__device__ __forceinline__ double illinois_point(const Illinois& s) { return s.b - s.fb * (s.b - s.a) / (s.fb - s.fa); }

__device__ __forceinline__ void illinois_update(Illinois& s, double c, double fc)
{
    if (s.done) return;
    const bool conv = fc == 0.0 || fabs(c - s.b) <= 1e-15;
    if (fc * s.fb < 0.0) {
        s.a = s.b;
        s.fa = s.fb;
    } else {
        s.fa *= 0.5;
    }
    s.b = c;
    s.fb = fc;
    s.done = conv;
}

}  // namespace cbc
