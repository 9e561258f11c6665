// cbc_linalg.h -- the in-register coupling solve shared by the lane and group
// mappings.
#pragma once

#include "cbc_device.h"

namespace cbc {

// Conditional exchanges of row k with rows from..n-1, in order: a row whose
// |a_ik| exceeds the current |a_kk| takes its place.
template <int n>
__device__ __forceinline__ void lu_cascade(double (&a)[n][n + 1], int k, int from)
{
#pragma unroll
    for (int i = from; i < n; ++i) {
        const bool sw = fabs(a[i][k]) > fabs(a[k][k]);
#pragma unroll
        for (int j = k; j <= n; ++j) {
            const double u = a[k][j];
            const double v = a[i][j];
            a[k][j] = sw ? v : u;
            a[i][j] = sw ? u : v;
        }
    }
}

// In-register LU with partial pivoting of the augmented n x (n+1) system.
// The pivot search is a sequence of conditional row exchanges so every index
// stays a compile-time constant (no scratch).  The first row holding the
// largest |a_ik| ends up as pivot, as in the CPU restatement.
//
// Exchanges are not rare in this model, but they follow a pattern: on the
// serving, mixed and plain ensembles (oracle LU, 69 000-1 300 000 solves per
// set) an exchange happens only in the first N - 1 (radial) columns, for 1-7 %
// of the solves per column, and with the next row in all but ~0.1 % of the
// solves (a 512-system serving-style set of 4 bubbles: 0.1 % at the second
// column take a row two down).  With 16-64 systems per wave the exchange code
// of those columns runs in a third to two thirds of the waves.  The rare far
// exchanges stay in registers too: solving those systems again from their
// staged rows, serially in LDS, made an N = 4 attempt 19.3 -> 27.8 us.  For n = 8 (the group mapping at N = 4), where some lane of the
// wave needs an exchange at a step, the cascade runs as its first exchange (row
// k + 1, n + 1 - k selects per row) and then the rest (rows k + 2.., tested
// against the diagonal after the first) only if some lane needs that: the same
// comparisons in the same order as one cascade, so every lane's result is the
// cascade's bit for bit (an N=4 attempt 21.2 -> 19.6 us).  For n <= 6 the
// second branch costs more than it saves (N=3 group 14.0 -> 14.3 us, lane N=2
// unchanged) and the cascade stays whole.  All branches are wave-uniform.
template <int n>
__device__ __forceinline__ void lu_solve(double (&a)[n][n + 1], double (&x)[n])
{
    constexpr bool kSplit = n >= 8;
    double inv[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
        bool need = false;
#pragma unroll
        for (int i = k + 1; i < n; ++i) need |= fabs(a[i][k]) > fabs(a[k][k]);
        if (k + 1 < n && __any(need)) {
            if constexpr (kSplit) {
                const bool next = fabs(a[k + 1][k]) > fabs(a[k][k]);
#pragma unroll
                for (int j = k; j <= n; ++j) {
                    const double u = a[k][j];
                    const double v = a[k + 1][j];
                    a[k][j] = next ? v : u;
                    a[k + 1][j] = next ? u : v;
                }
                bool far = false;
#pragma unroll
                for (int i = k + 2; i < n; ++i) far |= fabs(a[i][k]) > fabs(a[k][k]);
                if (__any(far)) lu_cascade<n>(a, k, k + 2);
            } else {
                lu_cascade<n>(a, k, k + 1);
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
