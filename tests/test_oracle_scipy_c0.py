"""The C oracle against the FP64 reference run: SciPy solve_ivp with DOP853 and with
Radau at rtol = atol = 1e-12 on the NumPy restatement of the model (the matrix form of
reference/CKM1D_reference.py, integrated as its solve_ivp call does, :848-856), on eight
systems of the 256-system plain workload over its 5 periods (fixture written by
tests/golden/make_c0_scipy.py).

* DOP853 (explicit 8th order) and Radau (implicit 5th order) agree with each other far
  inside the bar, so the fixture is the ODE's solution, not one method's;
* the C oracle at the workload's rtol = atol = 1e-9 is within the 1e-6 parity bar of it,
  and at 1e-12 within 1e-9: the per-unit C restatement and the Cash--Karp stepper
  integrate the same ODE as the matrix-form NumPy restatement;
* one fixture system is recomputed with DOP853 here, so the fixture cannot go stale.
"""
from pathlib import Path

import numpy as np

from coupledbubble_control import workloads
from oracle import ckm1d_numpy as NP
from oracle import ckm1d_oracle as O

FIX = Path(__file__).resolve().parent / "golden" / "c0_scipy_1e12.npz"


def _fixture():
    f = np.load(FIX)
    w = workloads.plain_cpu(seed=0)
    idx = f["systems"]
    np.testing.assert_array_equal(w.R0[idx], f["R0"])      # the workload generator is unchanged
    np.testing.assert_array_equal(w.X0[idx], f["X0"])
    np.testing.assert_array_equal(w.PA[:, idx], f["PA"])
    return w, f


def _as_state(y):
    """solve_ivp state rows (k, [R(2), x(2), U(2), V(2)]) -> (4, k, 2)."""
    return np.asarray(y).reshape(-1, 4, 2).transpose(1, 0, 2)


def rel_err(got, want):
    """Max over systems and quantities of |got - want| / (the quantity's max over the bubbles)."""
    scale = np.abs(want).max(axis=2, keepdims=True) + 1e-12
    return float((np.abs(got - want) / scale).max())


def test_dop853_and_radau_agree():
    _, f = _fixture()
    e = rel_err(_as_state(f["y_dop853"]), _as_state(f["y_radau"]))
    assert e <= 1e-8, e


def test_fixture_reproduces_one_system():
    w, f = _fixture()
    s = int(f["systems"][0])
    m = workloads.build_model(w.subset(np.array([s])))
    up, gp, dp, mx, y0 = NP.from_reference_layout(O.host_buffers_dict(m), 0, 2, 2)
    sol = NP.integrate(y0, 5.0, up, gp, dp, mx, kind="SW_N", method="DOP853", rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sol.y[:, -1], f["y_dop853"][0], rtol=1e-12, atol=1e-15)


def test_oracle_matches_scipy_reference():
    w, f = _fixture()
    idx = f["systems"]
    want = _as_state(f["y_dop853"])
    for tol, bar in ((1e-9, 1e-6), (1e-12, 1e-9)):
        sub = w.subset(idx)
        sub.rtol = sub.atol = tol
        m = workloads.build_model(sub)
        b = O.host_buffers_dict(m)
        O.solve(O.spec_from_model(m), b, kernel_steps=1 << 20, max_steps=1 << 26)
        assert np.all(b["status_flags"] == 1)
        e = rel_err(b["actual_state"][:, :, :2], want)
        print(f"oracle at rtol=atol={tol:g}: max rel err vs DOP853 1e-12 = {e:.2e}")
        assert e <= bar, (tol, e)
