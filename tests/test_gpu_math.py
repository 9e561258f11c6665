"""The kernels' FP64 elementary functions (csrc/cbc_math.h) evaluated on the device
through cbc_selftest_math and checked against numpy float64 (the host libm), in
units of the last place of the reference value.  Covers the full functions
(sincos, exp, log, pow, rcp) and the stage-reuse forms (fdlibm kernels without
reduction, the small-angle series, the addition theorem, the (1+e)^y series)
on the argument ranges the kernels use them on."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eval(name, x, y=None):
    import torch

    from coupledbubble_control.backends.hip import _abi

    dx = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    dy = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).cuda() if y is not None else None
    out = torch.empty(2 * dx.numel(), dtype=torch.float64, device="cuda")
    _abi.check(_abi.lib().cbc_selftest_math(_abi.MATH_CODES[name], dx.data_ptr(),
                                            dy.data_ptr() if dy is not None else None, out.data_ptr(),
                                            dx.numel(), None))
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(-1, 2)
    return o[:, 0], o[:, 1]


def _ulp_err(got, want, scale=None):
    scale = np.abs(want) if scale is None else scale
    return np.max(np.abs(got - want) / np.spacing(np.maximum(scale, 1e-300)))


def test_sincos_full_range():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-10, 10, 20000), rng.uniform(-1.6e6, 1.6e6, 20000),
                        np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 1e-300])])
    s, c = _eval("sincos", x)
    # absolute error in ulps of 1 (values near zero crossings carry the argument's rounding)
    assert np.max(np.abs(s - np.sin(x))) <= 8 * np.spacing(1.0)
    assert np.max(np.abs(c - np.cos(x))) <= 8 * np.spacing(1.0)
    small = np.abs(x) <= 10
    assert _ulp_err(s[small], np.sin(x[small]), np.maximum(np.abs(np.sin(x[small])), 0.25)) <= 2
    assert _ulp_err(c[small], np.cos(x[small]), np.maximum(np.abs(np.cos(x[small])), 0.25)) <= 2


def test_exp_log_pow_rcp():
    rng = np.random.default_rng(1)
    x = rng.uniform(-40, 40, 20000)
    assert _ulp_err(_eval("exp", x)[0], np.exp(x)) <= 2
    z = np.exp(rng.uniform(-30, 30, 20000))
    assert _ulp_err(_eval("log", z)[0], np.log(z), np.maximum(np.abs(np.log(z)), 1.0)) <= 2
    r = np.exp(rng.uniform(-5, 5, 20000))
    e = rng.uniform(-6, 6, 20000)
    got = _eval("pow", r, e)[0]
    # exp(y log x): the log's last-place error is scaled by |y log x|
    bound = (np.abs(e * np.log(r)) + 4.0) * np.spacing(1.0)
    assert np.all(np.abs(got / np.power(r, e) - 1.0) <= bound)
    d = rng.uniform(-1e3, 1e3, 20000)
    assert _ulp_err(_eval("rcp", d)[0], 1.0 / d) <= 1


def test_stage_reuse_forms():
    rng = np.random.default_rng(2)
    r = rng.uniform(-np.pi / 4, np.pi / 4, 20000)
    s, cm1 = _eval("sincos_kern", r)
    assert _ulp_err(s, np.sin(r), np.maximum(np.abs(r), 1e-3)) <= 2
    ref_cm1 = -2.0 * np.sin(r / 2) ** 2
    assert np.max(np.abs(cm1 - ref_cm1)) <= 4 * np.spacing(np.abs(ref_cm1).max())
    b = rng.uniform(-0.05, 0.05, 20000)
    s, cm1 = _eval("sincos_tiny", b)
    assert _ulp_err(s, np.sin(b)) <= 2
    ref_cm1 = -2.0 * np.sin(b / 2) ** 2
    assert _ulp_err(cm1, ref_cm1, np.maximum(np.abs(ref_cm1), 1e-20)) <= 4
    a = rng.uniform(-20, 20, 20000)
    s, c = _eval("rotate", a, b)
    assert np.max(np.abs(s - np.sin(a + b))) <= 8 * np.spacing(20.0)
    assert np.max(np.abs(c - np.cos(a + b))) <= 8 * np.spacing(20.0)
    e = rng.uniform(-0.01, 0.01, 20000)
    y = rng.uniform(-6, 6, 20000)
    got = _eval("pow1p", e, y)[0]
    want = np.exp(y * np.log1p(e))
    assert np.max(np.abs(got / want - 1.0)) <= 4e-16


def _solve(n, a):
    import torch

    from coupledbubble_control.backends.hip import _abi

    da = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    x = torch.empty((a.shape[0], n), dtype=torch.float64, device="cuda")
    _abi.check(_abi.lib().cbc_selftest_solve(n, da.data_ptr(), x.data_ptr(), a.shape[0], None))
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("n", [2, 4, 6, 8])
def test_coupling_solve_every_exchange_path(n):
    """The in-register LU on systems that need no exchange (column-diagonally dominant
    I + C: partial pivoting keeps the diagonal), an exchange with the next row only, an
    exchange with a row further down at nearly every step (the dominant system with its
    rows shuffled, so each column's pivot sits at a random row below), tiny diagonals
    (an elimination without the right pivots loses most digits there), a first pivot two
    or more rows down, and general systems; solutions against LAPACK's
    (numpy.linalg.solve)."""
    print(f"case: coupling solve n={n}", flush=True)
    rng = np.random.default_rng(70 + n)
    m = 512
    c = rng.uniform(-0.9, 0.9, (m, n, n)) / max(n - 1, 1)
    dom = np.concatenate([c, rng.uniform(-1, 1, (m, n, 1))], axis=2)
    dom[:, np.arange(n), np.arange(n)] = 1.0
    nxt = dom.copy()
    nxt[:, [0, 1]] = nxt[:, [1, 0]]
    shuf = np.stack([s[rng.permutation(n)] for s in dom])
    tiny = rng.uniform(-1, 1, (m, n, n + 1))
    tiny[:, np.arange(n), np.arange(n)] *= 1e-10
    gen = rng.uniform(-1, 1, (m, n, n + 1))
    cases = [("dominant", dom), ("next row", nxt), ("shuffled rows", shuf), ("tiny diagonal", tiny), ("general", gen)]
    if n > 2:
        # tiny diagonal and first sub-diagonal: the first pivot lies two or more rows
        # down; an elimination that exchanged with the next row at most is off by
        # 1e-4 to 5e-3 relative here (host check), the full rule by ~1e-13
        far = rng.uniform(-1, 1, (m, n, n + 1))
        far[:, np.arange(n), np.arange(n)] *= 1e-10
        far[:, np.arange(1, n), np.arange(n - 1)] *= 1e-10
        cases.append(("far pivot", far))
    for name, a in cases:
        A, b = a[:, :, :n], a[:, :, n]
        want = np.linalg.solve(A, b[..., None])[..., 0]
        got = _solve(n, a)
        cond = np.linalg.cond(A)
        err = np.abs(got - want).max(axis=1) / np.abs(want).max(axis=1)
        ok = cond < 1e8
        assert ok.sum() > m // 2, name
        assert np.all(err[ok] <= 1e-14 * n * cond[ok] + 1e-15), (name, float(err[ok].max()))
