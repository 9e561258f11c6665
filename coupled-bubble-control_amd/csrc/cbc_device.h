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
// (t, x) given the time harmonics sin/cos(th_k); ckm1d_def.py:40-155.
template <int NK, int FIELD>
__device__ __forceinline__ void field_at(const Drive<NK>& d, const double (&st)[NK], const double (&ct)[NK],
                                         double x, double& p, double& pt, double& px, double& ux)
{
    p = 0.0; pt = 0.0; px = 0.0; ux = 0.0;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        if constexpr (FIELD == kFieldConst) {
            p += d.A[k] * st[k];
            pt += d.Aw[k] * ct[k];
        } else {
            double sx, cx;
            sincos_fast(d.wx[k] * x + d.ph[k], &sx, &cx);
            if constexpr (FIELD == kFieldSwN) {
                p += d.A[k] * sx * st[k];
                pt += d.Aw[k] * sx * ct[k];
                px += d.Ak[k] * cx * st[k];
                ux += d.Au[k] * cx * ct[k];
            } else {
                p += d.A[k] * cx * st[k];
                pt += d.Aw[k] * cx * ct[k];
                px -= d.Ak[k] * sx * st[k];
                ux -= d.Au[k] * sx * ct[k];
            }
        }
    }
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

__device__ __forceinline__ void bubble_rhs(const Unit& up, double sp0, double sp3,
                                           double x0, double x2, double x3,
                                           double pa, double pat, double pax, double uax,
                                           BubbleOut& o)
{
    const double rx0 = rcp_fast(x0);
    const double p = pow_pos(x0, -sp0);  // (1/x0)^sp0
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

}  // namespace cbc
