/*
 * ckm1d_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C CPU restatement of the reference's coupled Keller--Miksis 1-D
 * (CKM1D) right-hand side and of its Cash--Karp RK45 ensemble stepper.  It is
 * the checker for the HIP kernels and the CPU baseline in bench.py; nothing in
 * the product links or loads it.  Only tests/, __graft_entry__.smoke() and the
 * bench.py cpu_baseline leg may use it.
 *
 * What it follows (paths relative to the reference repository root):
 *   RHS, per unit      backends/cupy_cuda/cuda_src/ckm1d_def.cuh:379-490
 *                      backends/numba/sysmtem_definitions/ckm1d_def.py:407-496
 *   acoustic field     ckm1d_def.py:40-155 (u_ac prefactor sp[4], no 2*pi)
 *   explicit coupling  ckm1d_def.cuh:66-178, reference/CKM1D_reference.py:526-580
 *   implicit coupling  ckm1d_def.cuh:180-376 (Richardson), CKM1D_reference.py:310-331
 *                      (matrix-free A*v); here solved either directly (LU with
 *                      partial pivoting, the canonical choice) or by the
 *                      reference's relaxed Richardson iteration.
 *   collision event    ckm1d_def.py:498-556 (threshold 0.1)
 *   stepper            backends/numba/rkck45_coupled_numba_stepper.py:116-706
 *                      backends/cupy_cuda/cuda_src/rkck45_coupled_cuda_sover.cu:206-846
 *   tableau            rkck45_coupled_cuda_sover.cu:137-147
 *   host launch loop   backends/numba/rkck45_coupled_numba_solver.py:28-92
 *
 * Pins (tests/test_oracle_*.py): Cash--Karp order conditions and polynomial
 * exactness; controller formulas on y'=lambda*y; equilibrium invariant,
 * textbook SI Keller--Miksis / translational equations, Minnaert resonance;
 * far-field limit, exchange symmetry; direct solve vs numpy.linalg.solve;
 * the NumPy matrix-form restatement (oracle/ckm1d_numpy.py) and SciPy
 * DOP853/Radau at rtol=atol=1e-12.
 *
 * Buffers use the reference layout (models/model.py:56-106):
 *   state (4, NS, TILE), unit_params (13, NS, TILE), system_params (5, NS),
 *   control_params (4*NK, NS), coupling (3, UPS, NS, TILE),
 *   dense_time (NDO, NS), dense_state (NDO, 4, NS, TILE).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the OpenMP loops over systems (the CPU baseline times 1 and all
 * cores); returns the count in effect (1 without OpenMP). */
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

#define OR_MAXN 32
#define OR_MAXK 8

enum { OR_FIELD_CONST = 0, OR_FIELD_SW_N = 1, OR_FIELD_SW_A = 2 };
enum { OR_LIN_DIRECT = 0, OR_LIN_RICHARDSON = 1 };

/* status codes: rkck45_coupled_numba_stepper.py:27-33 */
enum {
    OR_ACTIVE = 0, OR_SUCCESS = 1, OR_EVENT_TERMINAL = 2,
    OR_MAX_STEPS = -1, OR_STEP_SIZE_LIMIT = -2, OR_CONVERGENCE = -3, OR_NAN = -4
};

typedef struct {
    int n;                /* bubbles in the system (UPS) */
    int nk;               /* harmonic components (NK) */
    int field;            /* OR_FIELD_* */
    double up[13][OR_MAXN];
    double sp[5];
    double cp[4 * OR_MAXK];
    double mx[3][OR_MAXN][OR_MAXN];   /* mx[c][i][j]: effect of bubble j on bubble i */
    int linsolve;
    double lin_atol, lin_rtol, lin_relax;
    int lin_max_iter;
    double col_threshold; /* collision threshold, 0.1 in ckm1d_def.py:534 */
} or_system;

typedef struct {
    double atol, rtol, min_step, max_step, etol;
    int event_exact;      /* exact collision-event localisation (see or_event_root) */
} or_tol;

/* ---------------------------------------------------------------------- */
/* acoustic field: ckm1d_def.py:40-155                                     */
/* ---------------------------------------------------------------------- */
static void or_field(const or_system *S, double t, double x,
                     double *p, double *pt, double *px, double *ux)
{
    const double TWO_PI = 6.283185307179586476925286766559;
    const double w0 = TWO_PI * S->sp[1];
    const double l0 = TWO_PI * S->sp[2];
    const double u0 = S->sp[4];
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    for (int i = 0; i < S->nk; i++) {
        double amps = S->cp[i];
        double omega = S->cp[S->nk + i];
        double phase = S->cp[2 * S->nk + i];
        double wave = S->cp[3 * S->nk + i];
        double th = w0 * omega * t + phase;
        double st = sin(th), ct = cos(th);
        if (S->field == OR_FIELD_CONST) {
            a += amps * st;
            b += amps * omega * ct;
        } else {
            double xh = l0 * wave * x + phase;
            double sx = sin(xh), cx = cos(xh);
            if (S->field == OR_FIELD_SW_N) {
                a += amps * sx * st;
                b += amps * omega * sx * ct;
                c += amps * wave * cx * st;
                d += amps * u0 * cx * ct;
            } else {
                a += amps * cx * st;
                b += amps * omega * cx * ct;
                c -= amps * wave * sx * st;
                d -= amps * u0 * sx * ct;
            }
        }
    }
    *p = a; *pt = b; *px = c; *ux = d;
}

static double or_sign(double v) { return (double)((v > 0) - (v < 0)); }

/*
 * Dense LU solve with partial pivoting (the pivot row is the first row holding
 * the largest |a_ik|).  Back substitution multiplies by the reciprocal pivots
 * formed during elimination.
 */
static int or_lu_solve(int n, double *A /* n*n row-major */, double *b)
{
    double inv[4 * OR_MAXN];
    for (int k = 0; k < n; k++) {
        int piv = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(A[i * n + k]);
            if (v > best) { best = v; piv = i; }
        }
        if (best == 0.0) return -1;
        if (piv != k) {
            for (int j = 0; j < n; j++) {
                double tmp = A[k * n + j]; A[k * n + j] = A[piv * n + j]; A[piv * n + j] = tmp;
            }
            double tb = b[k]; b[k] = b[piv]; b[piv] = tb;
        }
        inv[k] = 1.0 / A[k * n + k];
        for (int i = k + 1; i < n; i++) {
            double f = A[i * n + k] * inv[k];
            for (int j = k + 1; j < n; j++) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[i * n + j] * b[j];
        b[i] = s * inv[i];
    }
    return 0;
}

/*
 * Right-hand side of one system.  y and dy are ordered [R(n), x(n), U(n), V(n)].
 * Returns the number of Richardson iterations used (0 for the direct solve),
 * or -1 when the Richardson iteration did not converge / the LU is singular.
 */
int or_rhs(const or_system *S, double t, const double *y, double *dy)
{
    const int n = S->n;
    const double *R = y, *X = y + n, *U = y + 2 * n, *V = y + 3 * n;
    double G[4][OR_MAXN], H[7][OR_MAXN];
    double b[2 * OR_MAXN];

    /* uncoupled part: ckm1d_def.py:418-468 */
    for (int i = 0; i < n; i++) {
        const double x0 = R[i], x1 = X[i], x2 = U[i], x3 = V[i];
        const double rx0 = 1.0 / x0;
        const double p = pow(rx0, S->sp[0]);
        double pa, pat, pax, uax;
        or_field(S, t, x1, &pa, &pat, &pax, &uax);
        const double up0 = S->up[0][i], up1 = S->up[1][i], up2 = S->up[2][i], up3 = S->up[3][i];
        const double up4 = S->up[4][i], up5 = S->up[5][i], up6 = S->up[6][i], up7 = S->up[7][i];
        const double up8 = S->up[8][i], up9 = S->up[9][i], up10 = S->up[10][i], up11 = S->up[11][i];
        const double N = (up0 + up1 * x2) * p
                       - up2 * (1 + up7 * x2)
                       - up3 * rx0
                       - up4 * x2 * rx0
                       - 1.5 * (1.0 - up7 * x2 * (1.0 / 3.0)) * x2 * x2
                       - (1 + up7 * x2) * up5 * pa
                       - up6 * pat * x0
                       + up8 * x3 * x3;
        const double D = x0 - up7 * x0 * x2 + up4 * up7;
        const double rD = 1.0 / D;
        const double Fb1 = -up10 * x0 * x0 * x0 * pax;
        const double Fd = -up11 * x0 * (x3 * S->sp[3] - uax);
        dy[i] = x2;
        dy[n + i] = x3;
        dy[2 * n + i] = N * rD;
        dy[3 * n + i] = 3 * (Fb1 + Fd) * up9 * rx0 * rx0 * rx0 - 3.0 * x2 * rx0 * x3;
        G[0][i] = rD;
        G[1][i] = x3 * rD;
        G[2][i] = x2 * rx0;
        G[3][i] = rx0 * rx0;
        const double x02 = x0 * x0, x03 = x02 * x0;
        H[0][i] = x0 * x2 * x2;
        H[1][i] = x02 * x2;
        H[2][i] = x02 * x2 * x3;
        H[3][i] = x03 * x3;
        H[4][i] = x03 * x3 * x3;
        H[5][i] = x02;
        H[6][i] = x03;
    }

    /* explicit coupling: ckm1d_def.py:163-246 */
    for (int i = 0; i < n; i++) {
        double rg0 = 0.0, rg1 = 0.0, tng = 0.0, tg2 = 0.0, tg3 = 0.0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            double delta = X[i] - X[j];
            double s = or_sign(delta);
            delta = fabs(delta);
            double rd = 1.0 / fmax(delta, 1e-30);
            double rd2 = rd * rd, srd2 = rd2 * s, rd3 = rd2 * rd;
            double m0 = S->mx[0][i][j], m1 = S->mx[1][i][j], m2 = S->mx[2][i][j];
            rg0 += m0 * (-2.0 * rd * H[0][j] - 2.5 * srd2 * H[2][j] - 1.0 * rd3 * H[4][j]);
            rg1 += m0 * (-0.5 * srd2 * H[1][j] - 0.5 * rd3 * H[3][j]);
            tng += m1 * (2.0 * srd2 * H[0][j] + 5.0 * rd3 * H[2][j]);
            tg2 += m1 * (srd2 * H[1][j] + rd3 * H[3][j]);
            tg3 += m2 * (srd2 * H[1][j] + rd3 * H[3][j]);
        }
        dy[2 * n + i] += rg0 * G[0][i] + rg1 * G[1][i];
        dy[3 * n + i] += tng + tg2 * G[2][i] + tg3 * G[3][i];
    }

    /* implicit coupling (I + C) v = b, b = [dU; dV]: ckm1d_def.cuh:180-376 */
    for (int i = 0; i < n; i++) { b[i] = dy[2 * n + i]; b[n + i] = dy[3 * n + i]; }

    if (S->linsolve == OR_LIN_DIRECT) {
        double A[4 * OR_MAXN * OR_MAXN];
        const int m = 2 * n;
        memset(A, 0, sizeof(double) * m * m);
        for (int i = 0; i < n; i++) {
            A[i * m + i] = 1.0;
            A[(n + i) * m + n + i] = 1.0;
            for (int j = 0; j < n; j++) {
                if (j == i) continue;
                double delta = X[i] - X[j];
                double s = or_sign(delta);
                double rd = 1.0 / fmax(fabs(delta), 1e-30);
                double rd2 = rd * rd, srd2 = rd2 * s, rd3 = rd2 * rd;
                double m0 = S->mx[0][i][j], m1 = S->mx[1][i][j];
                A[i * m + j] = G[0][i] * m0 * rd * H[5][j];
                A[i * m + n + j] = 0.5 * G[0][i] * m0 * srd2 * H[6][j];
                A[(n + i) * m + j] = -m1 * srd2 * H[5][j];
                A[(n + i) * m + n + j] = -m1 * rd3 * H[6][j];
            }
        }
        if (or_lu_solve(m, A, b) != 0) return -1;
        for (int i = 0; i < n; i++) { dy[2 * n + i] = b[i]; dy[3 * n + i] = b[n + i]; }
        return 0;
    }

    /* Richardson v <- v + w (b - A v) from v = b; ckm1d_def.py:302-404 */
    {
        double v[2 * OR_MAXN], r[2 * OR_MAXN];
        double bn2 = 0.0;
        for (int k = 0; k < 2 * n; k++) { v[k] = b[k]; bn2 += b[k] * b[k]; }
        double tol2 = fmax(S->lin_atol * S->lin_atol, S->lin_rtol * S->lin_rtol * bn2);
        int it;
        int conv = 0;
        for (it = 0; it < S->lin_max_iter; it++) {
            double rn2 = 0.0;
            for (int i = 0; i < n; i++) {
                double av_r = 0.0, av_t = 0.0;
                for (int j = 0; j < n; j++) {
                    if (j == i) continue;
                    double delta = X[i] - X[j];
                    double s = or_sign(delta);
                    double rd = 1.0 / fmax(fabs(delta), 1e-30);
                    double rd2 = rd * rd, srd2 = rd2 * s, rd3 = rd2 * rd;
                    double m0 = S->mx[0][i][j], m1 = S->mx[1][i][j];
                    av_r += m0 * (rd * H[5][j] * v[j] + 0.5 * srd2 * H[6][j] * v[n + j]);
                    av_t += -m1 * (srd2 * H[5][j] * v[j] + rd3 * H[6][j] * v[n + j]);
                }
                double ax_r = av_r * G[0][i] + v[i];
                double ax_t = av_t + v[n + i];
                r[i] = b[i] - ax_r;
                r[n + i] = b[n + i] - ax_t;
                rn2 += r[i] * r[i] + r[n + i] * r[n + i];
            }
            if (rn2 < tol2) { conv = 1; break; }
            for (int k = 0; k < 2 * n; k++) v[k] += S->lin_relax * r[k];
        }
        for (int i = 0; i < n; i++) { dy[2 * n + i] = v[i]; dy[3 * n + i] = v[n + i]; }
        return conv ? it : -1;
    }
}

/* collision gap: ckm1d_def.py:498-556; the chain order is the index order */
double or_event(const or_system *S, const double *y)
{
    const int n = S->n;
    double g = 1e30;
    for (int i = 0; i + 1 < n; i++) {
        double ri = y[i] * S->up[12][i];
        double rj = y[i + 1] * S->up[12][i + 1];
        double gap = fabs(y[n + i] - y[n + i + 1]) - (ri + rj) * (1.0 + S->col_threshold);
        if (gap < g) g = gap;
    }
    return g;
}

/*
 * Exact event mode: the root of the collision gap along one step y0 -> y1 of
 * length h.  R and x follow cubic Hermite interpolants of their end values
 * with slopes h*U and h*V (R' = U, x' = V in the dimensionless model); the gap
 * of the interpolated state is solved for by the Illinois method (modified
 * regula falsi, Dowell & Jarratt 1971) from the bracket (0, g0), (1, g1): at
 * most 60 iterations, stopping when the gap is 0 or two iterates differ by at
 * most 1e-15.  The reference's CPU path locates the same root with solve_ivp's
 * event search on its dense output (CKM1D_reference.py:702-714, 848-856).
 */
static double or_hermite3(double p0, double p1, double d0, double d1, double th)
{
    double t2 = th * th, t3 = t2 * th;
    return (2.0 * t3 - 3.0 * t2 + 1.0) * p0 + (t3 - 2.0 * t2 + th) * d0 + (3.0 * t2 - 2.0 * t3) * p1
           + (t3 - t2) * d1;
}

double or_event_root(const or_system *S, const double *y0, const double *y1, double h, double g0, double g1)
{
    const int n = S->n;
    double yc[2 * OR_MAXN];
    double a = 0.0, fa = g0, b = 1.0, fb = g1;
    for (int it = 0; it < 60; it++) {
        double c = b - fb * (b - a) / (fb - fa);
        for (int i = 0; i < n; i++) {
            yc[i] = or_hermite3(y0[i], y1[i], h * y0[2 * n + i], h * y1[2 * n + i], c);
            yc[n + i] = or_hermite3(y0[n + i], y1[n + i], h * y0[3 * n + i], h * y1[3 * n + i], c);
        }
        double fc = or_event(S, yc);
        int conv = (fc == 0.0) || fabs(c - b) <= 1e-15;
        if (fc * fb < 0.0) { a = b; fa = fb; } else { fa *= 0.5; }
        b = c; fb = fc;
        if (conv) break;
    }
    return b;
}


/* Cash--Karp tableau: rkck45_coupled_cuda_sover.cu:137-147 */
static const double CK_A[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
static const double CK_B[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
    {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
static const double CK_C[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
static const double CK_CP[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0,
                                277.0 / 14336.0, 1.0 / 4.0};

void or_tableau(double *a, double *b, double *c, double *cp)
{
    for (int i = 0; i < 6; i++) {
        a[i] = CK_A[i];
        c[i] = CK_C[i];
        cp[i] = CK_CP[i];
        for (int j = 0; j < 5; j++) b[i * 5 + j] = CK_B[i][j];
    }
}

/* Step-size controller factors: rkck45_coupled_numba_stepper.py:610-611 */
double or_fac_ok(double ratio) { return fmin(5.0, fmax(0.2, 0.9 * pow(ratio, 0.2))); }
double or_fac_bad(double ratio) { return fmin(0.5, fmax(0.1, 0.9 * pow(ratio, 0.25))); }

/* Per-system integrator state carried across launches. */
typedef struct {
    double t, t_end, t_begin, h, event;
    int flag;
    int64_t accepted, rejected, attempts;
    int dense_index;
} or_run;

/* Optional per-period statistics (used by the parameter-study workload). */
typedef struct {
    double period;        /* length of a statistics window in tau units (0 = off) */
    int first_window;     /* windows with index < first_window are not recorded */
    int n_windows;        /* capacity of max_r / mean_x per system */
    double *max_r;        /* [n_windows][n] */
    double *mean_x;       /* [n_windows][n] */
} or_period;

/*
 * One "launch" worth of adaptive steps for one system, following
 * rkck45_coupled_numba_stepper.py:179-706 lane by lane.
 * Dense output (NDO>0) and period windows are optional.
 */
static void or_store_dense(const or_system *S, or_run *st, const double *y, int ndo,
                           double *next_time, double min_dt,
                           double *dtime /* [ndo] */, double *dstate /* [ndo][4][n] */)
{
    if (ndo <= 0) return;
    int bi = st->dense_index;
    if (bi < ndo && *next_time <= st->t) {
        dtime[bi] = st->t;
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < S->n; i++) dstate[(bi * 4 + j) * S->n + i] = y[j * S->n + i];
        st->dense_index = bi + 1;
        *next_time = fmin(st->t + min_dt, st->t_end);
    }
}

static void or_window_update(const or_system *S, const or_period *P, double t0, double h,
                             const double *y0, const double *y1, double *wmax, double *wsum)
{
    (void)t0; (void)h; (void)P;
    const int n = S->n;
    for (int i = 0; i < n; i++) {
        if (y1[i] > wmax[i]) wmax[i] = y1[i];
        wsum[i] += 0.5 * h * (y0[n + i] + y1[n + i]);
    }
}

/* Leave the statistics of a window that is still open (budget exhausted or the
   system stopped) so that the next call can resume it. */
static void or_window_flush(const or_period *P, int wcur, double t, int n, const double *wmax, const double *wsum)
{
    if (!P || P->period <= 0.0) return;
    int widx = wcur - P->first_window;
    if ((t - wcur * P->period) > 1e-9 * P->period && widx >= 0 && widx < P->n_windows) {
        for (int i = 0; i < n; i++) {
            P->max_r[widx * n + i] = wmax[i];
            P->mean_x[widx * n + i] = wsum[i] / P->period;
        }
    }
}

int or_advance(const or_system *S, const or_tol *T, or_run *st, double *y, int kernel_steps,
               int ndo, double *dtime, double *dstate, const or_period *P)
{
    const int n = S->n;
    const int d = 4 * n;
    double k[6][4 * OR_MAXN], tmp[4 * OR_MAXN], ynew[4 * OR_MAXN];

    /* load: stepper.py:179-213 */
    int terminated = (st->t >= st->t_end);
    st->flag = terminated ? OR_SUCCESS : OR_ACTIVE;
    if (fabs(st->event) < T->etol) { terminated = 1; st->flag = OR_EVENT_TERMINAL; }
    if (terminated) return 0;

    double next_dense = st->t, min_dt = 0.0;
    if (ndo > 0) {
        min_dt = (ndo == 1) ? (st->t_end - st->t_begin) : (st->t_end - st->t_begin) / (ndo - 1);
        or_store_dense(S, st, y, ndo, &next_dense, min_dt, dtime, dstate);
    }
    /* period windows: a window boundary acts as an intermediate end time */
    double wmax[OR_MAXN], wsum[OR_MAXN];
    int wcur = -1;
    double wend = st->t_end;
    if (P && P->period > 0.0) {
        wcur = (int)floor(st->t / P->period + 1e-12);
        wend = fmin((wcur + 1) * P->period, st->t_end);
        int widx = wcur - P->first_window;
        int mid = (st->t - wcur * P->period) > 1e-9 * P->period;
        for (int i = 0; i < n; i++) {
            wmax[i] = y[i]; wsum[i] = 0.0;
            if (mid && widx >= 0 && widx < P->n_windows) {   /* resume a partial window */
                wmax[i] = P->max_r[widx * n + i];
                wsum[i] = P->mean_x[widx * n + i] * P->period;
            }
        }
    }

    double t = st->t, h = st->h;
    double g_old = or_event(S, y);

    int attempts = 0;
    for (int step = 0; step < kernel_steps; step++) {
        attempts++;
        st->attempts++;
        /* clamp: stepper.py:366-373 */
        if (h < T->min_step) h = T->min_step;
        if (h > T->max_step) h = T->max_step;
        double remain = wend - t;
        if (h > remain) h = remain;

        /* stages: stepper.py:379-480 */
        or_rhs(S, t, y, k[0]);
        for (int s = 1; s < 6; s++) {
            for (int c = 0; c < d; c++) {
                double acc = CK_B[s][0] * k[0][c];
                for (int q = 1; q < s; q++) acc += CK_B[s][q] * k[q][c];
                tmp[c] = y[c] + h * acc;
            }
            or_rhs(S, t + CK_A[s] * h, tmp, k[s]);
        }
        for (int c = 0; c < d; c++) {
            double acc = CK_C[0] * k[0][c];
            for (int q = 1; q < 6; q++) acc += CK_C[q] * k[q][c];
            ynew[c] = y[c] + h * acc;
        }
        double g_new = or_event(S, ynew);

        /* event handling: stepper.py:500-515; the exact mode accepts a crossing
           step within etol of the root and otherwise aims the retry at the root */
        int crossing = (g_old * g_new < 0.0);
        if (crossing && T->event_exact && fabs(g_new) < T->etol) crossing = 0;
        double aold = fabs(g_old);
        double ev_ratio = (aold / (aold + fabs(g_new) + 1e-15)) * 0.99;
        if (crossing && T->event_exact) ev_ratio = fmin(or_event_root(S, y, ynew, h, g_old, g_new), 0.99);

        int nan_state = 0;
        for (int c = 0; c < d; c++) if (!isfinite(ynew[c])) nan_state = 1;

        if (crossing) {
            /* no progress possible; a NaN step size (the retry factor of an infinite
               gap) passes every clamp, and the reference retries it until its step
               budget runs out: flagged as a non-finite step instead */
            if (!(h > T->min_step)) {
                st->flag = (h == h) ? OR_STEP_SIZE_LIMIT : OR_NAN;
                st->t = t; st->h = h; st->event = g_old;
                or_window_flush(P, wcur, t, n, wmax, wsum);
                return attempts;
            }
            h = h * ev_ratio;
            st->rejected++;
        } else {
            /* error norm: stepper.py:525-536 */
            double ratio = 1.0e30;
            int ok = 1;
            for (int c = 0; c < d; c++) {
                double acc = (CK_C[0] - CK_CP[0]) * k[0][c];
                for (int q = 1; q < 6; q++) acc += (CK_C[q] - CK_CP[q]) * k[q][c];
                double e = fabs(h * acc) + 1.0e-18;
                double tol = fmax(T->rtol * fmax(fabs(ynew[c]), fabs(y[c])), T->atol);
                double r = tol / e;
                if (r != r) nan_state = 1;
                if (r < ratio) ratio = r;
                if (r < 1.0) ok = 0;
            }
            if (nan_state) {
                st->flag = OR_NAN;
                st->t = t; st->h = h; st->event = g_old;
                or_window_flush(P, wcur, t, n, wmax, wsum);
                return attempts;
            }
            /* step size below the time resolution (the reference clamps at MIN_STEP
               and spins until its step budget runs out) */
            if (ok ? (t + h == t) : (h <= T->min_step)) {
                st->flag = OR_STEP_SIZE_LIMIT;
                st->t = t; st->h = h; st->event = g_old;
                or_window_flush(P, wcur, t, n, wmax, wsum);
                return attempts;
            }
            /* controller: stepper.py:585-618 */
            if (ok) {
                double h_acc = h;
                if (P && P->period > 0.0) or_window_update(S, P, t, h_acc, y, ynew, wmax, wsum);
                t = t + h_acc;
                memcpy(y, ynew, sizeof(double) * d);
                g_old = g_new;
                h = h_acc * or_fac_ok(ratio);
                st->accepted++;
            } else {
                h = h * or_fac_bad(ratio);
                st->rejected++;
            }
        }
        st->t = t;
        if (ndo > 0) or_store_dense(S, st, y, ndo, &next_dense, min_dt, dtime, dstate);

        /* window close */
        if (P && P->period > 0.0 && t >= wend && wend < st->t_end) {
            int widx = wcur - P->first_window;
            if (widx >= 0 && widx < P->n_windows) {
                for (int i = 0; i < n; i++) {
                    P->max_r[widx * n + i] = wmax[i];
                    P->mean_x[widx * n + i] = wsum[i] / P->period;
                }
            }
            wcur++;
            wend = fmin((wcur + 1) * P->period, st->t_end);
            for (int i = 0; i < n; i++) { wmax[i] = y[i]; wsum[i] = 0.0; }
        }

        /* termination: stepper.py:651-663 */
        if (t >= st->t_end) {
            if (P && P->period > 0.0) {
                int widx = wcur - P->first_window;
                if (widx >= 0 && widx < P->n_windows) {
                    for (int i = 0; i < n; i++) {
                        P->max_r[widx * n + i] = wmax[i];
                        P->mean_x[widx * n + i] = wsum[i] / P->period;
                    }
                }
            }
            st->flag = OR_SUCCESS;
            st->h = h; st->event = g_old;
            return attempts;
        }
        if (fabs(g_old) < T->etol) {
            st->flag = OR_EVENT_TERMINAL;
            st->h = h; st->event = g_old;
            or_window_flush(P, wcur, t, n, wmax, wsum);
            return attempts;
        }
    }
    st->h = h;
    st->event = g_old;
    or_window_flush(P, wcur, t, n, wmax, wsum);
    return attempts;
}

/* ---------------------------------------------------------------------- */
/* reference-layout driver                                                 */
/* ---------------------------------------------------------------------- */
typedef struct {
    int ns, ups, tile, nk, field, ndo;
    double atol, rtol, min_step, max_step, etol;
    int linsolve; double lin_atol, lin_rtol, lin_relax; int lin_max_iter;
    double col_threshold;
    int event_exact;
} or_config;

static void or_load_system(const or_config *C, int s, const double *unit_params,
                           const double *system_params, const double *control_params,
                           const double *coupling, or_system *S)
{
    const int NS = C->ns, TILE = C->tile, n = C->ups;
    S->n = n; S->nk = C->nk; S->field = C->field;
    for (int j = 0; j < 13; j++)
        for (int u = 0; u < n; u++) S->up[j][u] = unit_params[(size_t)j * NS * TILE + (size_t)s * TILE + u];
    for (int j = 0; j < 5; j++) S->sp[j] = system_params[(size_t)j * NS + s];
    for (int j = 0; j < 4 * C->nk; j++) S->cp[j] = control_params[(size_t)j * NS + s];
    /* GCM(c, j, s, i) = coupling[c][j][s][i] is the effect of j on i (ckm1d_def.cuh:56-57) */
    for (int c = 0; c < 3; c++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                S->mx[c][i][j] = coupling[(((size_t)c * n + j) * NS + s) * TILE + i];
    S->linsolve = C->linsolve;
    S->lin_atol = C->lin_atol; S->lin_rtol = C->lin_rtol;
    S->lin_relax = C->lin_relax; S->lin_max_iter = C->lin_max_iter;
    S->col_threshold = C->col_threshold;
}

/*
 * The host loop of rkck45_coupled_numba_solver.py:28-92 applied to every
 * system of a reference-layout ensemble: up to ceil(max_steps/kernel_steps)
 * launches of kernel_steps attempts; a system stops early when its flag is
 * set.  OpenMP over systems when compiled with -fopenmp.
 * total_steps / rejected_steps may be NULL.
 */
int or_solve_reference_layout(const or_config *C, int kernel_steps, int64_t max_steps,
                              double *actual_time, const double *time_end, const double *time_begin,
                              double *time_step, double *actual_state,
                              const double *unit_params, const double *system_params,
                              const double *control_params, const double *coupling,
                              int *dense_index, double *dense_time, double *dense_state,
                              double *actual_event, int *status_flags,
                              int64_t *total_steps, int64_t *rejected_steps)
{
    const int NS = C->ns, TILE = C->tile, n = C->ups, NDO = C->ndo;
    const int64_t max_launches = (max_steps + kernel_steps - 1) / kernel_steps;
    or_tol T = {C->atol, C->rtol, C->min_step, C->max_step, C->etol, C->event_exact};
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+:bad)
    for (int s = 0; s < NS; s++) {
        or_system S;
        or_load_system(C, s, unit_params, system_params, control_params, coupling, &S);
        double y[4 * OR_MAXN];
        for (int j = 0; j < 4; j++)
            for (int u = 0; u < n; u++) y[j * n + u] = actual_state[(size_t)j * NS * TILE + (size_t)s * TILE + u];
        or_run st;
        st.t = actual_time[s]; st.t_end = time_end[s]; st.t_begin = time_begin[s];
        st.h = time_step[s]; st.event = actual_event[s]; st.flag = 0;
        st.accepted = 0; st.rejected = 0; st.attempts = 0;
        st.dense_index = NDO > 0 ? dense_index[s] : 0;
        double *dt = NULL, *ds = NULL;
        if (NDO > 0) {
            dt = (double *)malloc(sizeof(double) * NDO);
            ds = (double *)malloc(sizeof(double) * NDO * 4 * n);
            for (int q = 0; q < NDO; q++) {
                dt[q] = dense_time[(size_t)q * NS + s];
                for (int j = 0; j < 4; j++)
                    for (int u = 0; u < n; u++)
                        ds[(q * 4 + j) * n + u] =
                            dense_state[(((size_t)q * 4 + j) * NS + s) * TILE + u];
            }
        }
        for (int64_t l = 0; l < max_launches; l++) {
            or_advance(&S, &T, &st, y, kernel_steps, NDO, dt, ds, NULL);
            if (st.flag != OR_ACTIVE) break;
        }
        for (int j = 0; j < 4; j++)
            for (int u = 0; u < n; u++) actual_state[(size_t)j * NS * TILE + (size_t)s * TILE + u] = y[j * n + u];
        actual_time[s] = st.t; time_step[s] = st.h; actual_event[s] = st.event;
        status_flags[s] = st.flag;
        if (total_steps) total_steps[s] = st.attempts;
        if (rejected_steps) rejected_steps[s] = st.rejected;
        if (NDO > 0) {
            dense_index[s] = st.dense_index;
            for (int q = 0; q < NDO; q++) {
                dense_time[(size_t)q * NS + s] = dt[q];
                for (int j = 0; j < 4; j++)
                    for (int u = 0; u < n; u++)
                        dense_state[(((size_t)q * 4 + j) * NS + s) * TILE + u] = ds[(q * 4 + j) * n + u];
            }
            free(dt); free(ds);
        }
        if (st.flag < 0) bad++;
    }
    return bad;
}

/*
 * Period-window variant for the parameter-study workload: integrates every
 * system to time_end with window boundaries every `period` tau, recording the
 * maximum R/R0 and the trapezoidal mean x of each window with index >=
 * first_window into max_r / mean_x shaped (n_windows, NS, TILE).
 */
int or_solve_periodic(const or_config *C, int64_t max_attempts, double period, int first_window,
                      int n_windows,
                      double *actual_time, const double *time_end, double *time_step,
                      double *actual_state, const double *unit_params, const double *system_params,
                      const double *control_params, const double *coupling,
                      double *actual_event, int *status_flags,
                      int64_t *total_steps, int64_t *rejected_steps,
                      double *max_r, double *mean_x, int s_begin, int s_end)
{
    const int NS = C->ns, TILE = C->tile, n = C->ups;
    or_tol T = {C->atol, C->rtol, C->min_step, C->max_step, C->etol, C->event_exact};
    int bad = 0;
    if (s_end > NS) s_end = NS;
#pragma omp parallel for schedule(dynamic, 1) reduction(+:bad)
    for (int s = s_begin; s < s_end; s++) {
        or_system S;
        or_load_system(C, s, unit_params, system_params, control_params, coupling, &S);
        double y[4 * OR_MAXN];
        for (int j = 0; j < 4; j++)
            for (int u = 0; u < n; u++) y[j * n + u] = actual_state[(size_t)j * NS * TILE + (size_t)s * TILE + u];
        or_run st;
        st.t = actual_time[s]; st.t_end = time_end[s]; st.t_begin = 0.0;
        st.h = time_step[s]; st.event = actual_event[s]; st.flag = 0;
        st.accepted = 0; st.rejected = 0; st.attempts = 0; st.dense_index = 0;
        double *mr = (double *)malloc(sizeof(double) * (size_t)(n_windows > 0 ? n_windows : 1) * n);
        double *mx = (double *)malloc(sizeof(double) * (size_t)(n_windows > 0 ? n_windows : 1) * n);
        for (int q = 0; q < n_windows; q++)       /* caller's arrays: resumes open windows */
            for (int u = 0; u < n; u++) {
                mr[q * n + u] = max_r[((size_t)q * NS + s) * TILE + u];
                mx[q * n + u] = mean_x[((size_t)q * NS + s) * TILE + u];
            }
        or_period P = {period, first_window, n_windows, mr, mx};
        /* windows force a restart of the attempt loop at each boundary */
        int64_t used = 0;
        while (used < max_attempts) {
            int chunk = (int)((max_attempts - used) > 1000000 ? 1000000 : (max_attempts - used));
            used += or_advance(&S, &T, &st, y, chunk, 0, NULL, NULL, &P);
            if (st.flag != OR_ACTIVE) break;
        }
        for (int j = 0; j < 4; j++)
            for (int u = 0; u < n; u++) actual_state[(size_t)j * NS * TILE + (size_t)s * TILE + u] = y[j * n + u];
        actual_time[s] = st.t; time_step[s] = st.h; actual_event[s] = st.event;
        status_flags[s] = st.flag;
        if (total_steps) total_steps[s] = st.attempts;
        if (rejected_steps) rejected_steps[s] = st.rejected;
        for (int q = 0; q < n_windows; q++)
            for (int u = 0; u < n; u++) {
                max_r[((size_t)q * NS + s) * TILE + u] = mr[q * n + u];
                mean_x[((size_t)q * NS + s) * TILE + u] = mx[q * n + u];
            }
        free(mr); free(mx);
        if (st.flag < 0) bad++;
    }
    return bad;
}

/* Single-system RHS in the reference layout, for unit tests. */
int or_rhs_reference_layout(const or_config *C, int s, double t, const double *unit_params,
                            const double *system_params, const double *control_params,
                            const double *coupling, const double *y, double *dy)
{
    or_system S;
    or_load_system(C, s, unit_params, system_params, control_params, coupling, &S);
    return or_rhs(&S, t, y, dy);
}

double or_event_reference_layout(const or_config *C, int s, const double *unit_params,
                                 const double *system_params, const double *control_params,
                                 const double *coupling, const double *y)
{
    or_system S;
    or_load_system(C, s, unit_params, system_params, control_params, coupling, &S);
    return or_event(&S, y);
}

/* Generic RK driver on y' = lambda*y for controller tests. */
int or_linear_trace(double lambda, double y0, double t_end, double h0, const or_tol *T,
                    int max_attempts, double *t_out, double *h_out, int *ok_out)
{
    double t = 0.0, y = y0, h = h0;
    int a;
    for (a = 0; a < max_attempts; a++) {
        if (h < T->min_step) h = T->min_step;
        if (h > T->max_step) h = T->max_step;
        if (h > t_end - t) h = t_end - t;
        double k[6];
        k[0] = lambda * y;
        for (int s = 1; s < 6; s++) {
            double acc = CK_B[s][0] * k[0];
            for (int q = 1; q < s; q++) acc += CK_B[s][q] * k[q];
            k[s] = lambda * (y + h * acc);
        }
        double acc = CK_C[0] * k[0], eacc = (CK_C[0] - CK_CP[0]) * k[0];
        for (int q = 1; q < 6; q++) { acc += CK_C[q] * k[q]; eacc += (CK_C[q] - CK_CP[q]) * k[q]; }
        double ynew = y + h * acc;
        double e = fabs(h * eacc) + 1.0e-18;
        double tol = fmax(T->rtol * fmax(fabs(ynew), fabs(y)), T->atol);
        double r = tol / e;
        int ok = r >= 1.0;
        t_out[a] = t; h_out[a] = h; ok_out[a] = ok;
        if (ok) { t += h; y = ynew; h = h * or_fac_ok(r); }
        else h = h * or_fac_bad(r);
        if (t >= t_end) { a++; break; }
    }
    return a;
}
