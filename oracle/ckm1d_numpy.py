"""NumPy / SciPy restatement of the reference CPU model.  TEST INFRASTRUCTURE ONLY.

Follows coupledbubble_control/reference/CKM1D_reference.py:
  * parameters              MultiBubble._update_control_parameters  :775-829
  * acoustic field          _PA/_PAT/_GRADP/_UAC                    :49-304
  * uncoupled RHS           _ode_function                           :583-624
  * explicit coupling       _explicit_coupling (matrix form)        :526-580
  * implicit coupling       A = I + C assembled as in the commented
                            block :672-693 and solved exactly with
                            numpy.linalg.solve (the reference iterates
                            BiCG on the matrix-free _AxV, :310-411)
  * collision event         _collision_event                        :702-714
but with the collision threshold 0.1 of the GPU kernels
(backends/numba/sysmtem_definitions/ckm1d_def.py:534; the CPU script uses
0.25, :39) and with material properties passed in.  Integration uses
scipy.integrate.solve_ivp (DOP853 / Radau) at tight tolerances; this is the
FP64 trajectory oracle, independent of the Cash--Karp step sequence.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Mat:
    PV: float = 0.0
    RHO: float = 998.0
    ST: float = 0.0725
    VIS: float = 0.001
    CL: float = 1500.0
    P0: float = 1.013e5
    PE: float = 1.4


def water_25c() -> Mat:
    return Mat(PV=3.17e3, RHO=998.0, ST=0.072, VIS=0.89e-3, CL=1497.0, P0=1.0e5, PE=1.4)


def parameters(R0_um, FREQ_khz, PA_bar, PS_rad, ref_freq_khz, mat: Mat):
    """(up (13, n), gp (5,), dp (4k,), mx (3, n, n)) as in CKM1D_reference.py:775-829,
    with up[10] = 4 pi/3 R0^3 and up[11] = 12 pi mu R0 (= the GPU forms 2*2pi/3, 6*2pi)."""
    R0 = np.asarray(R0_um, dtype=np.float64) * 1e-6
    FREQ = np.asarray(FREQ_khz, dtype=np.float64) * 1e3
    PA = np.asarray(PA_bar, dtype=np.float64) * 1e5
    PS = np.asarray(PS_rad, dtype=np.float64)
    k = FREQ.size
    n = R0.size
    PV, RHO, ST, VIS, CL, P0, PE = mat.PV, mat.RHO, mat.ST, mat.VIS, mat.CL, mat.P0, mat.PE
    wr = 2.0 * np.pi * ref_freq_khz * 1e3
    lr = CL / (ref_freq_khz * 1e3)
    up = np.zeros((13, n))
    up[0] = (2.0 * ST / R0 + P0 - PV) * (2.0 * np.pi / R0 / wr) ** 2.0 / RHO
    up[1] = (1.0 - 3.0 * PE) * (2 * ST / R0 + P0 - PV) * (2.0 * np.pi / R0 / wr) / CL / RHO
    up[2] = (P0 - PV) * (2.0 * np.pi / R0 / wr) ** 2.0 / RHO
    up[3] = (2.0 * ST / R0 / RHO) * (2.0 * np.pi / R0 / wr) ** 2.0
    up[4] = 4.0 * VIS / RHO / (R0 ** 2.0) * (2.0 * np.pi / wr)
    up[5] = ((2.0 * np.pi / R0 / wr) ** 2.0) / RHO
    up[6] = ((2.0 * np.pi / wr) ** 2.0) / CL / RHO / R0
    up[7] = R0 * wr / (2 * np.pi) / CL
    up[8] = (0.5 * lr / R0) ** 2
    up[9] = (2.0 * np.pi) / RHO / R0 / lr / (wr * R0) ** 2.0
    up[10] = 4 * np.pi / 3 * R0 ** 3
    up[11] = 12 * np.pi * VIS * R0
    up[12] = R0 / lr
    mx = np.zeros((3, n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                mx[0, i, j] = R0[j] ** 3 / R0[i] ** 2 / lr
                mx[1, i, j] = 3.0 * (R0[j] / lr) ** 3
                mx[2, i, j] = (18 * VIS / RHO) * (2 * np.pi / wr) * (R0[j] / lr) ** 3 / (R0[i] ** 2)
    gp = np.array([3 * PE, 1.0 / wr, lr / (2 * np.pi), CL, 1.0 / RHO / CL])
    dp = np.zeros(4 * k)
    dp[:k] = PA
    dp[k:2 * k] = 2.0 * np.pi * FREQ
    dp[2 * k:3 * k] = PS
    dp[3 * k:] = 2.0 * np.pi * FREQ / CL
    return up, gp, dp, mx


def field(t, x, gp, dp, kind: str):
    """(p, dp/dt, dp/dx, u_ac) at positions x; CKM1D_reference.py:49-304."""
    k = dp.size // 4
    p = np.zeros_like(x)
    pt = np.zeros_like(x)
    px = np.zeros_like(x)
    ux = np.zeros_like(x)
    for i in range(k):
        th = 2 * np.pi * gp[1] * dp[k + i] * t + dp[2 * k + i]
        if kind == "CONST":
            p += dp[i] * np.sin(th)
            pt += dp[i] * dp[k + i] * np.cos(th)
            continue
        xh = 2 * np.pi * gp[2] * dp[3 * k + i] * x + dp[2 * k + i]
        if kind == "SW_N":
            p += dp[i] * np.sin(xh) * np.sin(th)
            pt += dp[i] * dp[k + i] * np.sin(xh) * np.cos(th)
            px += dp[i] * dp[3 * k + i] * np.cos(xh) * np.sin(th)
            ux += dp[i] * gp[4] * np.cos(xh) * np.cos(th)
        elif kind == "SW_A":
            p += dp[i] * np.cos(xh) * np.sin(th)
            pt += dp[i] * dp[k + i] * np.cos(xh) * np.cos(th)
            px -= dp[i] * dp[3 * k + i] * np.sin(xh) * np.sin(th)
            ux -= dp[i] * gp[4] * np.sin(xh) * np.cos(th)
        else:
            raise ValueError(kind)
    return p, pt, px, ux


def coupling_parts(x, up, gp, dp, mx, t, kind):
    """Uncoupled derivative, G (4, n), H (7, n), r_delta and sign matrices."""
    n = x.size // 4
    x0, x1, x2, x3 = x[:n], x[n:2 * n], x[2 * n:3 * n], x[3 * n:]
    rx0 = 1.0 / x0
    pw = rx0 ** gp[0]
    pa, pat, pax, uax = field(t, x1, gp, dp, kind)
    N = ((up[0] + up[1] * x2) * pw - up[2] * (1 + up[7] * x2) - up[3] * rx0 - up[4] * x2 * rx0
         - 1.5 * (1.0 - up[7] * x2 * (1.0 / 3.0)) * x2 * x2 - (1 + up[7] * x2) * up[5] * pa
         - up[6] * pat * x0 + up[8] * x3 * x3)
    D = x0 - up[7] * x0 * x2 + up[4] * up[7]
    rD = 1.0 / D
    Fb1 = -up[10] * x0 ** 3 * pax
    Fd = -up[11] * x0 * (x3 * gp[3] - uax)
    du = 3.0 * ((Fb1 + Fd) * up[9] * rx0 * rx0 - x2 * x3) * rx0
    dx = np.concatenate([x2, x3, N * rD, du])
    delta = x1[:, None] - x1[None, :]
    s = np.sign(delta)
    np.fill_diagonal(delta, 1.0)
    r_delta = 1.0 / np.abs(delta)
    np.fill_diagonal(r_delta, 0.0)
    G = np.stack([rD, x3 * rD, x2 * rx0, rx0 ** 2])
    H = np.stack([x0 * x2 ** 2, x0 ** 2 * x2, x0 ** 2 * x2 * x3, x0 ** 3 * x3, x0 ** 3 * x3 ** 2, x0 ** 2,
                  x0 ** 3])
    return dx, G, H, r_delta, s


def explicit_coupling(r_delta, s, mx, G, H):
    """Right-hand-side correction; CKM1D_reference.py:526-580."""
    n = G.shape[1]
    dx_c = np.zeros(2 * n)
    cm = mx[0] * r_delta
    dx_c[:n] += -2.0 * (cm @ H[0]) * G[0]
    cm = cm * r_delta
    dx_c[:n] += -0.5 * ((cm * s) @ H[1]) * G[1] - 2.5 * ((cm * s) @ H[2]) * G[0]
    cm = cm * r_delta
    dx_c[:n] += -0.5 * (cm @ H[3]) * G[1] - 1.0 * (cm @ H[4]) * G[0]
    cm = mx[1] * r_delta ** 2
    dx_c[n:] += 2.0 * ((cm * s) @ H[0]) + ((cm * s) @ H[1]) * G[2]
    cm = cm * r_delta
    dx_c[n:] += 5.0 * (cm @ H[2]) + (cm @ H[3]) * G[2]
    cm = mx[2] * r_delta ** 2
    dx_c[n:] += ((cm * s) @ H[1]) * G[3]
    cm = cm * r_delta
    dx_c[n:] += (cm @ H[3]) * G[3]
    return dx_c


def implicit_matrix(r_delta, s, mx, G, H):
    """A = I + C (2n x 2n); the assembly of CKM1D_reference.py:672-693."""
    n = G.shape[1]
    A = np.zeros((2 * n, 2 * n))
    cm = mx[0] * r_delta
    A[:n, :n] = G[0][:, None] * (cm * H[5][None, :])
    cm2 = cm * r_delta * s
    A[:n, n:] = 0.5 * G[0][:, None] * (cm2 * H[6][None, :])
    c10 = mx[1] * r_delta ** 2
    A[n:, :n] = -(c10 * s) * H[5][None, :]
    A[n:, n:] = -(c10 * r_delta) * H[6][None, :]
    A += np.eye(2 * n)
    return A


def matvec(r_delta, s, mx, G, H, v):
    """Matrix-free A v; CKM1D_reference.py:310-331."""
    n = G.shape[1]
    x = np.zeros_like(v)
    C00 = mx[0] * r_delta
    x[:n] += G[0] * (C00 @ (H[5] * v[:n])) + v[:n]
    C01 = C00 * r_delta * s
    x[:n] += 0.5 * G[0] * (C01 @ (H[6] * v[n:]))
    C10 = mx[1] * r_delta ** 2
    x[n:] += -(C10 * s) @ (H[5] * v[:n])
    C11 = C10 * r_delta
    x[n:] += -C11 @ (H[6] * v[n:]) + v[n:]
    return x


def rhs(t, x, up, gp, dp, mx, kind="SW_N"):
    """Full coupled RHS with the exact coupling solve."""
    n = x.size // 4
    dx, G, H, r_delta, s = coupling_parts(x, up, gp, dp, mx, t, kind)
    dx[2 * n:] += explicit_coupling(r_delta, s, mx, G, H)
    if n > 1:
        A = implicit_matrix(r_delta, s, mx, G, H)
        dx[2 * n:] = np.linalg.solve(A, dx[2 * n:])
    return dx


def collision_event(x, up, threshold=0.1):
    n = x.size // 4
    r = x[:n] * up[12]
    x1 = x[n:2 * n]
    if n < 2:
        return 1e30
    return float(np.min(np.abs(x1[:-1] - x1[1:]) - (r[:-1] + r[1:]) * (1 + threshold)))


def integrate(y0, t_end, up, gp, dp, mx, kind="SW_N", method="DOP853", rtol=1e-12, atol=1e-12, t0=0.0,
              t_eval=None, threshold=0.1):
    """solve_ivp integration with the terminal collision event; returns the OdeResult."""
    from scipy.integrate import solve_ivp

    def ev(t, y):
        return collision_event(y, up, threshold)

    ev.terminal = True
    ev.direction = 0
    return solve_ivp(lambda t, y: rhs(t, y, up, gp, dp, mx, kind), (t0, t_end), np.asarray(y0, float),
                     method=method, rtol=rtol, atol=atol, events=ev, t_eval=t_eval, dense_output=False)


def from_reference_layout(bufs: dict, s: int, n: int, nk: int):
    """(up, gp, dp, mx, y0) of system s from reference-layout buffers."""
    up = bufs["unit_params"][:, s, :n].copy()
    gp = bufs["system_params"][:, s].copy()
    dp = bufs["control_params"][:, s].copy()
    cm = bufs["coupling_matrices"]  # (3, n, NS, TILE): [c][j][s][i] effect of j on i
    mx = np.transpose(cm[:, :, s, :n], (0, 2, 1)).copy()
    y0 = bufs["actual_state"][:, s, :n].reshape(-1).copy()
    return up, gp, dp, mx, y0
