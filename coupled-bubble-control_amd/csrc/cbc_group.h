// cbc_group.h -- group mapping: TILE = next_pow2(N) lanes per CKM1D system,
// one lane per bubble, 64/TILE systems per wave (one wave per workgroup), for
// N = 2..4; for N = 5..16 two lanes per bubble (16 lanes per system for N = 5..8,
// one system per wave for N = 9..16).
//
// This is the reference's thread-per-bubble layout (rkck45_coupled_cuda_sover.cu:228-239)
// with a different coupling solve: instead of the block-synchronous Richardson
// sweep (ckm1d_def.cuh:252-376), for N = 5..16 lane i of a group owns row i
// (radial) and lane TILE + i row N + i (translational) of the 2N x (2N+1)
// augmented system in VGPRs, both carrying bubble i, and the group runs
// Gauss--Jordan with partial pivoting: per column the pivot is the row with
// the largest |a_ik| among rows not yet used (ties: lowest row index), found by one
// max across the group over keys that pack |a|, the row and the sign (entries
// closer than 2^-45 relative count as ties); its owner's row reaches the other
// lanes by ds_bpermute (for N = 9..16, one system per wave, by v_readlane) and
// every lane eliminates that column from its row.  For N = 2..4 lane i builds
// both rows of bubble i, the group stages its n rows in LDS once (pairs: by DPP)
// and every lane solves the whole system in registers with the lane mapping's LU
// (identical results on every lane, one sync instead of n).
// Bubble data needed by the coupling sums (x, H0..H6) and the
// coupling constants mx are staged in LDS per system; pairs (N = 2) pass the
// bubble data and the rows by DPP broadcasts instead.  All lanes of a group
// reduce the error ratio and the collision gap to identical values, so the
// step controller runs redundantly and stays consistent without a broadcast.
#pragma once

#include <type_traits>
#include <utility>

#include "cbc_device.h"
#include "cbc_kernels.h"
#include "cbc_linalg.h"

namespace cbc {

constexpr int pow2_ceil(int v) { return v <= 1 ? 1 : 2 * pow2_ceil((v + 1) / 2); }

template <int N>
struct GroupShape {
    static constexpr int T = group_tile(N);   // lanes per bubble set: the buffers' TILE
    // N = 5..16: two lanes per bubble, lane i (i < T) with its radial coupling row
    // and lane T + i with its translational row; both hold the bubble's state and
    // step it identically, and only the first stores it.  N = 9..16 (`wave`): one
    // system per wave, lanes 32 + i and 48 + i carry bubble i once more (same
    // state, no part in the solve), so the pivot's owner is one lane of the wave
    static constexpr bool split = group_lanes(N) != T;
    static constexpr bool wave = group_lanes(N) == 64;
    static constexpr int G = group_lanes(N);  // lanes per system
    static constexpr int S = 64 / G;          // systems per wave
    static constexpr int n = 2 * N;           // coupling unknowns
    static constexpr bool redundant = !split;  // every lane solves the whole system (N <= 4)
};

template <int N>
struct GroupLds {
    static constexpr int G = GroupShape<N>::T;
    static constexpr int n = 2 * N;
    static constexpr bool R = GroupShape<N>::redundant;
    // Most frequently accessed first: ds_read2/ds_write2 address two doubles at
    // 8-bit offsets in units of 8 bytes (up to 2040 bytes from the base
    // register), so the staged rows and the bubble data stay in reach of the
    // group's base address without an address add per access.
    double sol[R ? 1 : n];          // solved accelerations (Gauss--Jordan)
    double rows[R ? n : 1][n + 1];  // the augmented system (redundant solve)
    double xh[8][G];           // x, H0..H6 of every bubble
    double ev[2][G];           // collision radius and position
    double mx[3][N][N];        // mx[c][i][j]: effect of bubble j on bubble i
};

// Reductions inside a group of G <= 16 lanes (every lane gets the result; the
// two lanes of a bubble in a 32-lane group hold the same bubble values, so its
// per-bubble reductions run over each 16-lane half),
// by DPP row permutations instead of LDS-routed shuffles: partner lane i ^ 1
// and i ^ 2 (quad_perm), then 7 - i within 8 lanes (row_half_mirror) and
// 15 - i within 16 (row_mirror).  After the quad steps every lane of a quad
// holds the quad's result, so each mirror step merges two whole sub-groups.
// min, OR and max do not depend on the order in which elements meet, so the result is a butterfly's (xor partners through
// LDS-routed shuffles, as these reductions were first written) bit for bit.  The
// reductions and broadcasts run in wave-uniform code only (all lanes active).
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;

// A plain DPP move (no "old" operand tied to the destination, so the source
// is not copied first: two VALU moves per double saved at every step).  Every
// source lane of these permutations lies in the lane's own row of 16 and every
// lane is active where they run, so no lane takes the undefined fallback value.
template <int Ctl>
__device__ __forceinline__ int dpp_int(int v)
{
    return __builtin_amdgcn_mov_dpp(v, Ctl, 0xF, 0xF, false);
}

template <int Ctl>
__device__ __forceinline__ double dpp_double(double v)
{
    const int lo = dpp_int<Ctl>(__double2loint(v));
    const int hi = dpp_int<Ctl>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

template <int G, typename Step>
__device__ __forceinline__ void group_steps(Step step)
{
    static_assert(G >= 1 && G <= 16 && (G & (G - 1)) == 0, "group of 1..16 lanes, a power of two");
    if constexpr (G >= 2) step(std::integral_constant<int, kDppXor1>{});
    if constexpr (G >= 4) step(std::integral_constant<int, kDppXor2>{});
    if constexpr (G >= 8) step(std::integral_constant<int, kDppHalfMirror>{});
    if constexpr (G >= 16) step(std::integral_constant<int, kDppMirror>{});
}

// Minimum of two doubles, neither a signalling NaN (as max_quiet below).
__device__ __forceinline__ double min_quiet(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <int G>
__device__ __forceinline__ double group_min(double v)
{
    v = __builtin_canonicalize(v);
    group_steps<G>([&](auto c) { v = min_quiet(v, dpp_double<decltype(c)::value>(v)); });
    return v;
}

template <int G>
__device__ __forceinline__ bool group_any(bool v)
{
    int x = v ? 1 : 0;
    group_steps<G>([&](auto c) { x |= dpp_int<decltype(c)::value>(x); });
    return x != 0;
}

// Pivot key of a Gauss--Jordan candidate: |a| with its 7 low bits replaced by
// (63 - row) << 1 | sign(a).  The largest key of the group is the largest |a|
// (ties within 2^-45 relative go to the lower row), the lowest row among equal
// ones, and it carries the winner's row, so one max over the group replaces the
// (value, lowest row) argmax.  NaN and infinite entries give NaN keys, which the
// max ignores; used rows carry -1, whose key decodes to row 63, owned by no lane.
__device__ __forceinline__ double pivot_key(double a, int row)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(a));
    const unsigned long long low = ((unsigned long long)(63 - row) << 1) | (__builtin_signbit(a) ? 1ull : 0ull);
    return __longlong_as_double((long long)((b & ~0x7Full) | low));
}

__device__ __forceinline__ int pivot_key_row(double key)
{
    return 63 - (int)(((unsigned long long)__double_as_longlong(key) & 0x7Full) >> 1);
}

// Maximum of two doubles, neither a signalling NaN, as one v_max_f64 (IEEE
// mode: a quiet NaN operand yields the other operand, as fmax).  fmax of an
// operand the compiler cannot prove quiet -- a DPP or permlane result -- first
// quiets it with a maximum of the operand with itself, a second FP64 operation
// on every step of a reduction; the group reductions quiet their value once and
// every later operand is a maximum (minimum) or a moved one, so the results are
// those of fmax (fmin) bit for bit.  The compiler's hazard check counts the statement as a VALU
// write and puts the wait states a DPP, permlane or readlane read of the result
// needs after it, as after its own v_max_f64.
__device__ __forceinline__ double max_quiet(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <int G>
__device__ __forceinline__ double group_max(double v)
{
    v = __builtin_canonicalize(v);
    group_steps<G>([&](auto c) { v = max_quiet(v, dpp_double<decltype(c)::value>(v)); });
    return v;
}

// Maximum over a group of G lanes; a 32-lane group takes the 16-lane maximum of
// each row, then exchanges its two rows by v_permlane16_swap (with both operands
// the same value, its two results are the even row's and the odd row's value of
// the lane's column).
template <int G>
__device__ __forceinline__ double group_max_wide(double v)
{
    if constexpr (G == 32) {
        v = group_max<16>(v);
        const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
        return max_quiet(__hiloint2double(hi[0], lo[0]), __hiloint2double(hi[1], lo[1]));
    } else {
        return group_max<G>(v);
    }
}

// Value of lane j of the lane's group, for groups of 2 or 4 lanes (inside one
// quad): a DPP quad_perm broadcast, no LDS.
template <int G, int J>
__device__ __forceinline__ double group_bcast(double v)
{
    static_assert(G == 2 || G == 4, "quad broadcast serves groups of 2 or 4 lanes");
    constexpr int ctl = G == 4 ? J * 0x55 : (J | J << 2 | (2 + J) << 4 | (2 + J) << 6);
    return dpp_double<ctl>(v);
}

// Orders the group's LDS stores before the loads that follow them in the other
// lanes of the group.  A workgroup is one wave here (64 threads) and a wave's LDS
// instructions execute in issue order, so wavefront-scope fences are the whole
// synchronisation: unlike the workgroup barrier they emit no lgkmcnt(0) wait,
// so the loads issue right behind the stores instead of after their completion.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of v in the lane at byte address `addr` (4 x lane) for every lane: two
// ds_bpermute (the LDS crossbar; every lane of the wave active).
__device__ __forceinline__ double bperm_double(int addr, double v)
{
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Value of v in lane `lane` (wave-uniform) for every lane.
__device__ __forceinline__ double readlane_double(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

template <int... J, typename F>
__device__ __forceinline__ void for_lanes(std::integer_sequence<int, J...>, F f)
{
    (f(std::integral_constant<int, J>{}), ...);
}

// Per-lane constants of one bubble and its system.
template <int NK>
struct GroupParams {
    Unit up;
    double rcol;
    double sp0, sp3;
    Drive<NK> d;
};

// Harmonics and polytropic factor of the lane's bubble at one stage.
template <int NK>
struct GroupStage {
    double st[NK], ct[NK];
    double sx[NK], cx[NK];
    double pf;
};

// First stage of an attempt, in full; rxb = 1/(R/R0).
template <int NK, int FIELD>
__device__ __forceinline__ void group_stage_base(const GroupParams<NK>& P, double t, const double (&y)[4],
                                                 GroupStage<NK>& B, double& rxb)
{
    time_base<NK>(P.d, t, B.st, B.ct);
    if constexpr (FIELD != kFieldConst) space_base<NK>(P.d, y[1], B.sx, B.cx);
    B.pf = pow_pos(y[0], -P.sp0);
    rxb = rcp_fast(y[0]);
}

// A later stage at (tq, x) from the first stage B at (t, y); `vote` marks
// lanes that carry a live bubble.
template <int NK, int FIELD>
__device__ __forceinline__ void group_stage_next(const GroupParams<NK>& P, const GroupStage<NK>& B, double rxb,
                                                 double t, const double (&y)[4], double tq, const double (&x)[4],
                                                 bool vote, GroupStage<NK>& S)
{
    time_stage<NK>(P.d, B.st, B.ct, t, tq, vote, S.st, S.ct);
    if constexpr (FIELD != kFieldConst) {
        if (vote && space_far<NK>(P.d, y[1], x[1])) space_base<NK>(P.d, x[1], S.sx, S.cx);
        else space_shift<NK>(P.d, B.sx, B.cx, y[1], x[1], S.sx, S.cx);
    }
    if (vote && pf_far(P.sp0, y[0], rxb, x[0])) S.pf = pow_pos(x[0], -P.sp0);
    else S.pf = pf_shift(P.sp0, B.pf, y[0], rxb, x[0]);
}

// Coupled RHS for the lane's bubble i.  y, dy = (R, x, U, V) of bubble i.
template <int N, int NK, int FIELD>
__device__ __forceinline__ void group_rhs(const GroupParams<NK>& P, GroupLds<N>& L, int i, int half, bool vu,
                                          const GroupStage<NK>& S, const double (&y)[4], double (&dy)[4])
{
    constexpr int n = 2 * N;
    double pa, pat, pax, uax;
    field_from<NK, FIELD>(P.d, S.st, S.ct, S.sx, S.cx, pa, pat, pax, uax);
    BubbleOut o;
    bubble_rhs(P.up, P.sp3, y[0], y[2], y[3], pa, pat, pax, uax, S.pf, o);
    dy[0] = y[2];
    dy[1] = y[3];
    if constexpr (N == 1) {
        dy[2] = o.dU;
        dy[3] = o.dV;
        return;
    } else {
        // x, H0..H6 of every bubble of the group.  Pairs (N = 2) exchange them,
        // and below the coupling rows, by DPP broadcasts inside the quad: the
        // same values without an LDS round trip (an attempt 11.1 -> 10.3 us).
        // Groups of 4 keep LDS: broadcasting their 72-entry system costs more
        // than the LDS trip (N = 4 19.0 -> 20.8 us with both exchanges by DPP,
        // 19.5 with this one alone).
        constexpr int G = GroupShape<N>::G;
        double xh[8][N];
        if constexpr (G == 2) {
            const double own[8] = {y[1], o.H0, o.H1, o.H2, o.H3, o.H4, o.H5, o.H6};
            for_lanes(std::make_integer_sequence<int, N>{}, [&](auto jc) {
                constexpr int j = decltype(jc)::value;
#pragma unroll
                for (int c = 0; c < 8; ++c) xh[c][j] = group_bcast<G, j>(own[c]);
            });
            // no LDS below for pairs (mx is written before the barrier that
            // follows every system load), so no barrier either
        } else {
            if (vu && half == 0) {
                L.xh[0][i] = y[1];
                L.xh[1][i] = o.H0;
                L.xh[2][i] = o.H1;
                L.xh[3][i] = o.H2;
                L.xh[4][i] = o.H3;
                L.xh[5][i] = o.H4;
                L.xh[6][i] = o.H5;
                L.xh[7][i] = o.H6;
            }
            wave_sync();
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int j = 0; j < N; ++j) xh[c][j] = L.xh[c][j];
        }

        // explicit coupling sums and the two matrix rows owned by this lane (split
        // groups: the lane's one row in ar, the radial row for half 0, else the
        // translational one)
        double ar[n + 1], at[n + 1];
        double rg0 = 0.0, rg1 = 0.0, tng = 0.0, tg2 = 0.0, tg3 = 0.0;
        const int ii = vu ? i : 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double xj = xh[0][j];
            const double delta = y[1] - xj;
            const double s = sgn(delta);
            const double r = rcp_fast(fmax(fabs(delta), 1e-30));
            const double r2 = r * r;
            const double sr2 = r2 * s;
            const double r3 = r2 * r;
            const double m0 = L.mx[0][ii][j], m1 = L.mx[1][ii][j], m2 = L.mx[2][ii][j];
            const double H0 = xh[1][j], H1 = xh[2][j], H2 = xh[3][j], H3 = xh[4][j], H4 = xh[5][j];
            const double H5 = xh[6][j], H6 = xh[7][j];
            const bool self = (j == i);
            if (!self) {
                rg0 += m0 * (-2.0 * r * H0 - 2.5 * sr2 * H2 - 1.0 * r3 * H4);
                rg1 += m0 * (-0.5 * sr2 * H1 - 0.5 * r3 * H3);
                tng += m1 * (2.0 * sr2 * H0 + 5.0 * r3 * H2);
                tg2 += m1 * (sr2 * H1 + r3 * H3);
                tg3 += m2 * (sr2 * H1 + r3 * H3);
            }
            if constexpr (GroupShape<N>::split) {
                const bool tr = (half & 1) != 0;   // translational-row lane
                const double cu = tr ? -m1 * sr2 : o.G0 * m0 * r;
                const double cv = tr ? -m1 * r3 : 0.5 * o.G0 * m0 * sr2;
                ar[j] = self ? (tr ? 0.0 : 1.0) : cu * H5;
                ar[N + j] = self ? (tr ? 1.0 : 0.0) : cv * H6;
            } else {
                ar[j] = self ? 1.0 : o.G0 * m0 * r * H5;
                ar[N + j] = self ? 0.0 : 0.5 * o.G0 * m0 * sr2 * H6;
                at[j] = self ? 0.0 : -m1 * sr2 * H5;
                at[N + j] = self ? 1.0 : -m1 * r3 * H6;
            }
        }
        if constexpr (GroupShape<N>::split) {
            ar[n] = (half & 1) ? o.dV + (tng + tg2 * o.G2 + tg3 * o.G3) : o.dU + (rg0 * o.G0 + rg1 * o.G1);
        } else {
            ar[n] = o.dU + (rg0 * o.G0 + rg1 * o.G1);
            at[n] = o.dV + (tng + tg2 * o.G2 + tg3 * o.G3);
        }

        if constexpr (GroupShape<N>::redundant) {
            // stage the n rows, then every lane solves the whole system in VGPRs
            double a[n][n + 1];
            if constexpr (G == 2) {
                for_lanes(std::make_integer_sequence<int, N>{}, [&](auto jc) {
                    constexpr int r = decltype(jc)::value;
#pragma unroll
                    for (int j = 0; j <= n; ++j) {
                        a[r][j] = group_bcast<G, r>(ar[j]);
                        a[N + r][j] = group_bcast<G, r>(at[j]);
                    }
                });
            } else {
                if (vu) {
#pragma unroll
                    for (int j = 0; j <= n; ++j) {
                        L.rows[i][j] = ar[j];
                        L.rows[N + i][j] = at[j];
                    }
                }
                wave_sync();
#pragma unroll
                for (int r = 0; r < n; ++r)
#pragma unroll
                    for (int j = 0; j <= n; ++j) a[r][j] = L.rows[r][j];
            }
            double v[n];
            lu_solve<n>(a, v);
            double d2 = v[0], d3 = v[N];
#pragma unroll
            for (int j = 1; j < N; ++j) {
                d2 = (j == ii) ? v[j] : d2;
                d3 = (j == ii) ? v[N + j] : d3;
            }
            dy[2] = d2;
            dy[3] = d3;
        } else if constexpr (GroupShape<N>::wave) {
        // Gauss--Jordan with partial pivoting, one row per lane, one system per wave
        // (N = 9..16): as below, but the pivot's owner is the same lane for the whole
        // wave, so its 1 / a_kk and row reach every lane by v_readlane (through
        // SGPRs) instead of an LDS round trip and a sync per column; the arithmetic
        // is the same.  Lanes 32..63 are no candidates (key -1).
        const bool cand = vu && half < 2;
        const int me = (half & 1) ? N + i : i;
        if (cand) L.sol[me] = __builtin_nan("");
        bool used = !cand;
        int col = 0;
        double inv_me = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const double key = used ? -1.0 : pivot_key(ar[k], me);
            // every candidate lane holds the same maximum; lane 0 is one of them
            const int row = __builtin_amdgcn_readfirstlane(pivot_key_row(group_max_wide<32>(key)));
            const int src = row < N ? row : (row < 2 * N ? GroupShape<N>::T + row - N : 0);  // owner lane (63: none)
            const bool own = cand && row == me;
            const double inv = readlane_double(rcp_fast(ar[k]), src);
            inv_me = own ? inv : inv_me;
            if (own) {
                used = true;
                col = k;
            }
            const double f = own ? 0.0 : ar[k] * inv;
#pragma unroll
            for (int j = k + 1; j <= n; ++j) ar[j] -= f * readlane_double(ar[j], src);
        }
        if (cand) L.sol[col] = ar[n] * inv_me;
        wave_sync();
        dy[2] = L.sol[ii];
        dy[3] = L.sol[N + ii];
        } else {
        // Gauss--Jordan with partial pivoting, one row per lane (N = 5..8, 16
        // lanes): per column one max over the pivot keys of the group (DPP steps
        // within the 16-lane row); the owner's exact 1 / a_kk and row reach the
        // group's lanes by ds_bpermute from the owner's lane (the LDS crossbar, no
        // LDS stores, sync or loads), and every lane eliminates the column from its
        // row.  Every solution slot starts as NaN (the sync before this solve
        // ordered it after the previous solve's reads): a column that finds no
        // finite unused row (NaN entries) has no owner, so some slot is never
        // written and the accelerations come out NaN, never stale values of the
        // previous solve.
        const int me = half ? N + i : i;
        if (vu) L.sol[me] = __builtin_nan("");
        bool used = !vu;
        int col = 0;
        double inv_me = 0.0;
        const int gbase = ((int)threadIdx.x / GroupShape<N>::G) * GroupShape<N>::G;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            const double key = used ? -1.0 : pivot_key(ar[k], me);
            const int row = pivot_key_row(group_max_wide<GroupShape<N>::G>(key));
            const bool own = vu && row == me;
            // byte address of the owner's lane for ds_bpermute (row 63: none)
            const int addr = 4 * (gbase + (row < N ? row : (row < 2 * N ? GroupShape<N>::T + row - N : 0)));
            const double inv = bperm_double(addr, rcp_fast(ar[k]));
            inv_me = own ? inv : inv_me;
            if (own) {
                used = true;
                col = k;
            }
            const double f = own ? 0.0 : ar[k] * inv;
#pragma unroll
            for (int j = k + 1; j <= n; ++j) ar[j] -= f * bperm_double(addr, ar[j]);
        }
        if (vu) L.sol[col] = ar[n] * inv_me;
        wave_sync();
        dy[2] = L.sol[ii];
        dy[3] = L.sol[N + ii];
        }
    }
}

// Collision gap of the system, identical on every lane of the group.
template <int N>
__device__ __forceinline__ double group_event(GroupLds<N>& L, int i, bool wr, bool vu, double rcol,
                                              const double (&y)[4], double colf)
{
    if constexpr (N == 1) {
        return 1e30;
    } else {
        if (vu && wr) {
            L.ev[0][i] = y[0] * rcol;
            L.ev[1][i] = y[1];
        }
        wave_sync();
        double gap = 1e30;
        if (vu && i + 1 < N) {
            const double ri = L.ev[0][i], rj = L.ev[0][i + 1];
            gap = fabs(L.ev[1][i] - L.ev[1][i + 1]) - (ri + rj) * colf;
        }
        wave_sync();  // ev is rewritten by the next call
        return group_min<GroupShape<N>::T>(gap);
    }
}

// Root theta of the gap along the step y0 -> y1 of length h (exact event mode),
// identical on every lane of the group.  Every lane of the wave runs all
// iterations (the gap evaluation syncs and reduces across the wave); a group that converged keeps
// its iterate, so the result is the early-exit one of the lane mapping.
template <int N>
__device__ double group_event_root(GroupLds<N>& L, int i, bool wr, bool vu, double rcol, const double (&y0)[4],
                                   const double (&y1)[4], double h, double colf, double g0, double g1)
{
    Illinois st{0.0, g0, 1.0, g1, false};
    for (int it = 0; it < kIllinoisIters; ++it) {
        const double c = illinois_point(st);
        double yc[4];
        yc[0] = hermite3(y0[0], y1[0], h * y0[2], h * y1[2], c);
        yc[1] = hermite3(y0[1], y1[1], h * y0[3], h * y1[3], c);
        yc[2] = 0.0;
        yc[3] = 0.0;
        const double g = group_event<N>(L, i, wr, vu, rcol, yc, colf);
        illinois_update(st, c, g);
    }
    return st.b;
}

}  // namespace cbc
