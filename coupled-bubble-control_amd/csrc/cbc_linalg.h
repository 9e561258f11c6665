// cbc_linalg.h -- the in-register coupling solve shared by the lane and group
// mappings.
#pragma once

#include "cbc_device.h"

namespace cbc {

// In-register LU with partial pivoting of the augmented n x (n+1) system.
// The pivot search is a sequence of conditional row exchanges so every index
// stays a compile-time constant (no scratch).  The first row holding the
// largest |a_ik| ends up as pivot, as in the CPU restatement.  The exchange
// code of a step runs only when some lane of the wave needs an exchange there
// (wave-uniform branch).  Exchanges are not rare in this model: in a wave of
// 16 systems of 4 bubbles some system needs one in most evaluations, so an
// exchange path that is cheap only when seldom taken (a serial solve of the
// staged rows) made the group kernel 3.6x slower, while the cascade costs ~20 %
// of an N = 4 attempt.
template <int n>
__device__ __forceinline__ void lu_solve(double (&a)[n][n + 1], double (&x)[n])
{
    double inv[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
        bool need = false;
#pragma unroll
        for (int i = k + 1; i < n; ++i) need |= fabs(a[i][k]) > fabs(a[k][k]);
        if (__any(need))
#pragma unroll
        for (int i = k + 1; i < n; ++i) {
            const bool sw = fabs(a[i][k]) > fabs(a[k][k]);
#pragma unroll
            for (int j = k; j <= n; ++j) {
                const double u = a[k][j];
                const double v = a[i][j];
                a[k][j] = sw ? v : u;
                a[i][j] = sw ? u : v;
            }
        }
        inv[k] = rcp_fast(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i][k] * inv[k];
#pragma unroll
            for (int j = k + 1; j <= n; ++j) a[i][j] -= f * a[k][j];
        }
    }
#pragma unroll
    for (int i = n - 1; i >= 0; --i) {
        double s = a[i][n];
#pragma unroll
        for (int j = i + 1; j < n; ++j) s -= a[i][j] * x[j];
        x[i] = s * inv[i];
    }
}

}  // namespace cbc
