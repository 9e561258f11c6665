/*
 * oracle_driver.c -- runs the C oracle (oracle/ckm1d_oracle.c, compiled into this
 * executable) on reference-layout buffers read from a file, for the sanitizer test
 * (tests/test_sanitizers_cpu.py builds it with -fsanitize=address,undefined).
 *
 * Input file: int32 header[8] = {mode (0 host loop, 1 period windows), ns, ups,
 * tile, nk, field, kernel_steps, n_windows}, float64 params[8] = {atol, rtol,
 * min_step, max_step, etol, col_threshold, period, first_window}, then float64
 * arrays actual_time (ns), time_end (ns), time_begin (ns), time_step (ns),
 * actual_state (4 ns tile), unit_params (13 ns tile), system_params (5 ns),
 * control_params (4 nk ns), coupling (3 ups ns tile), actual_event (ns).
 * Output file: actual_state, actual_time, time_step, actual_event (float64),
 * status_flags (int32), total_steps, rejected_steps (int64), then for mode 1
 * max_r and mean_x (float64, n_windows ns tile).
 */
#include <stdio.h>

#include "../../oracle/ckm1d_oracle.c"

static double *rd(FILE *f, size_t n)
{
    double *p = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!p || fread(p, sizeof(double), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
    return p;
}

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in out\n", argv[0]); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror("input"); return 2; }
    int32_t h[8];
    double q[8];
    if (fread(h, sizeof(int32_t), 8, f) != 8 || fread(q, sizeof(double), 8, f) != 8) { fprintf(stderr, "header\n"); return 2; }
    const int mode = h[0], ns = h[1], ups = h[2], tile = h[3], nk = h[4], field = h[5], ks = h[6], nw = h[7];
    const size_t nst = (size_t)ns * tile;
    double *t = rd(f, ns), *te = rd(f, ns), *tb = rd(f, ns), *hs = rd(f, ns);
    double *st = rd(f, 4 * nst), *up = rd(f, 13 * nst), *sp = rd(f, 5 * (size_t)ns);
    double *cp = rd(f, 4 * (size_t)nk * ns), *cm = rd(f, 3 * (size_t)ups * nst), *ev = rd(f, ns);
    fclose(f);
    or_config C;
    memset(&C, 0, sizeof(C));
    C.ns = ns; C.ups = ups; C.tile = tile; C.nk = nk; C.field = field; C.ndo = 0;
    C.atol = q[0]; C.rtol = q[1]; C.min_step = q[2]; C.max_step = q[3]; C.etol = q[4];
    C.linsolve = OR_LIN_DIRECT; C.lin_atol = 1e-9; C.lin_rtol = 1e-9; C.lin_relax = 1.0; C.lin_max_iter = 100;
    C.col_threshold = q[5]; C.event_exact = 0;
    int32_t *flags = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
    int64_t *tot = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
    int64_t *rej = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
    double *mr = NULL, *mx = NULL;
    if (mode == 0) {
        or_solve_reference_layout(&C, ks, (int64_t)1 << 26, t, te, tb, hs, st, up, sp, cp, cm, NULL, NULL, NULL, ev,
                                  flags, tot, rej);
    } else {
        mr = (double *)calloc((size_t)(nw ? nw : 1) * nst, sizeof(double));
        mx = (double *)calloc((size_t)(nw ? nw : 1) * nst, sizeof(double));
        or_solve_periodic(&C, (int64_t)1 << 26, q[6], (int)q[7], nw, t, te, hs, st, up, sp, cp, cm, ev, flags,
                          tot, rej, mr, mx, 0, ns);
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) { perror("output"); return 2; }
    fwrite(st, sizeof(double), 4 * nst, o);
    fwrite(t, sizeof(double), ns, o);
    fwrite(hs, sizeof(double), ns, o);
    fwrite(ev, sizeof(double), ns, o);
    fwrite(flags, sizeof(int32_t), ns, o);
    fwrite(tot, sizeof(int64_t), ns, o);
    fwrite(rej, sizeof(int64_t), ns, o);
    if (mode == 1) {
        fwrite(mr, sizeof(double), (size_t)nw * nst, o);
        fwrite(mx, sizeof(double), (size_t)nw * nst, o);
    }
    fclose(o);
    free(t); free(te); free(tb); free(hs); free(st); free(up); free(sp); free(cp); free(cm); free(ev);
    free(flags); free(tot); free(rej); free(mr); free(mx);
    printf("ok\n");
    return 0;
}
