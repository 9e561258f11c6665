"""Pins of the Cash--Karp tableau and the step controller of the oracle.

The oracle's coefficients (oracle/ckm1d_oracle.c, restating
rkck45_coupled_cuda_sover.cu:137-147) are checked against what the
mathematics fixes, not against a retyped table: row sums, quadrature and
order conditions, exactness on polynomials, the linear stability polynomial
derived independently from the Butcher matrix, and the observed convergence
order.  The controller (rkck45_coupled_numba_stepper.py:525-618) is checked
against closed-form factors and against a step trace recomputed from the
stability polynomials.
"""
import math

import numpy as np
import pytest

from oracle import ckm1d_oracle as O


@pytest.fixture(scope="module")
def tab():
    a, b, c, cp = O.tableau()
    A = np.zeros((6, 6))
    A[:, :5] = b
    return a, A, c, cp


def test_row_sums_equal_nodes(tab):
    a, A, c, cp = tab
    np.testing.assert_allclose(A.sum(axis=1), a, atol=1e-15)


def test_weights_sum_to_one_and_error_weights_to_zero(tab):
    a, A, c, cp = tab
    assert abs(c.sum() - 1.0) < 1e-15
    assert abs(cp.sum() - 1.0) < 1e-15
    assert abs((c - cp).sum()) < 1e-15


@pytest.mark.parametrize("order,weights", [(5, "c"), (4, "cp")])
def test_order_conditions(tab, order, weights):
    """All Butcher order conditions up to the method order (trees of order <= 5 for c,
    <= 4 for c*), written in matrix form b^T Phi(tree) = 1/gamma(tree)."""
    a, A, c, cp = tab
    w = c if weights == "c" else cp
    e = np.ones(6)
    C = a
    conds = {
        1: [(e, 1.0)],
        2: [(C, 1 / 2)],
        3: [(C ** 2, 1 / 3), (A @ C, 1 / 6)],
        4: [(C ** 3, 1 / 4), (C * (A @ C), 1 / 8), (A @ C ** 2, 1 / 12), (A @ A @ C, 1 / 24)],
        5: [(C ** 4, 1 / 5), (C ** 2 * (A @ C), 1 / 10), (C * (A @ C ** 2), 1 / 15), (C * (A @ A @ C), 1 / 30),
            ((A @ C) ** 2, 1 / 20), (A @ C ** 3, 1 / 20), (A @ (C * (A @ C)), 1 / 40), (A @ A @ C ** 2, 1 / 60),
            (A @ A @ A @ C, 1 / 120)],
    }
    for q in range(1, order + 1):
        for phi, target in conds[q]:
            assert abs(w @ phi - target) < 1e-14, (q, w @ phi, target)
    if weights == "cp":
        # the embedded solution is exactly 4th order: some 5th-order condition fails
        assert any(abs(w @ phi - t) > 1e-6 for phi, t in conds[5])


def _one_step(f, t, y, h, tab):
    a, A, c, cp = tab
    k = []
    for i in range(6):
        yi = y + h * sum(A[i, j] * k[j] for j in range(i))
        k.append(f(t + a[i] * h, yi))
    k = np.array(k)
    return y + h * (c @ k), h * ((c - cp) @ k)


def test_polynomial_exactness(tab):
    """y' = t^4 is integrated exactly by the 5th-order weights; the error estimate
    vanishes for integrands of degree <= 3."""
    y5, err = _one_step(lambda t, y: t ** 4, 0.3, 0.0, 0.7, tab)
    exact = (1.0 ** 5 - 0.3 ** 5) / 5
    assert abs(y5 - exact) < 1e-15
    for deg in range(4):
        _, err = _one_step(lambda t, y, d=deg: t ** d, 0.1, 0.0, 0.5, tab)
        assert abs(err) < 1e-15


def test_convergence_order(tab):
    """Local error of the 5th-order solution on y' = -y + sin(t) y^2 scales as h^6."""
    from scipy.integrate import solve_ivp

    f = lambda t, y: -y + np.sin(t) * y * y  # noqa: E731
    y0, t0 = 0.7, 0.2
    errs = []
    hs = [0.08, 0.04, 0.02]
    for h in hs:
        ex = solve_ivp(f, (t0, t0 + h), [y0], method="DOP853", rtol=3e-14, atol=1e-18).y[0, -1]
        y5, _ = _one_step(f, t0, y0, h, tab)
        errs.append(abs(y5 - ex))
    slopes = [math.log(errs[i] / errs[i + 1], 2) for i in range(2)]
    assert all(5.5 < s < 6.5 for s in slopes), slopes


def test_controller_closed_forms():
    """fac_ok = min(5, max(0.2, 0.9 r^0.2)), fac_bad = min(0.5, max(0.1, 0.9 r^0.25))."""
    assert O.fac_ok(1.0) == pytest.approx(0.9, rel=1e-15)
    assert O.fac_ok(1e12) == 5.0
    assert O.fac_ok(32.0) == pytest.approx(0.9 * 2.0, rel=1e-15)      # 32^0.2 = 2
    assert O.fac_bad(1.0) == 0.5
    assert O.fac_bad(1e-12) == 0.1
    assert O.fac_bad(1.0 / 16.0) == pytest.approx(0.9 * 0.5, rel=1e-15)  # (1/16)^0.25 = 1/2
    assert O.fac_ok(0.5 ** 5 / 0.9 ** 5 * 0.2 ** 5) == pytest.approx(0.2, rel=1e-12)


def test_linear_trace_follows_stability_polynomials(tab):
    """On y' = lam y every stage is linear, so y_new = P(z) y and err = E(z) y with
    P(z) = 1 + z c^T (I - zA)^-1 1 and E(z) = z (c - c*)^T (I - zA)^-1 1.  Replaying the
    controller with these polynomials must reproduce the oracle's attempt sequence."""
    a, A, c, cp = tab
    lam, y0, t_end, h0, atol, rtol = -3.0, 1.0, 2.0, 1e-6, 1e-9, 1e-9
    t_o, h_o, ok_o = O.linear_trace(lam, y0, t_end, h0, atol, rtol)
    t, y = 0.0, y0
    I6 = np.eye(6)
    e = np.ones(6)
    assert len(t_o) > 10
    for i in range(len(t_o)):
        h = h_o[i]                     # replay from the oracle's attempt (rounding does not accumulate)
        assert t == pytest.approx(t_o[i], rel=1e-13, abs=1e-15)
        z = lam * h
        v = np.linalg.solve(I6 - z * A, e)
        P = 1.0 + z * (c @ v)
        E = z * ((c - cp) @ v)
        ynew = P * y
        est = abs(E * y) + 1e-18
        tol = max(rtol * max(abs(ynew), abs(y)), atol)
        r = tol / est
        ok = r >= 1.0
        assert ok == ok_o[i]
        if ok:
            t += h
            y = ynew
            h_next = h * min(5.0, max(0.2, 0.9 * r ** 0.2))
        else:
            h_next = h * min(0.5, max(0.1, 0.9 * r ** 0.25))
        if i + 1 < len(t_o):
            # the error estimate is a difference of O(1) stage sums, so it carries
            # ~1e-9 relative rounding; the clamp to the remaining interval is exact
            want = min(max(h_next, 1e-16), 0.1, t_end - t)
            assert h_o[i + 1] == pytest.approx(want, rel=1e-6)
    assert t == pytest.approx(t_end, abs=1e-12)
    assert abs(y - y0 * math.exp(lam * t_end)) < 1e-8
