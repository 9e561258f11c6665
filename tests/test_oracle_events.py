"""Collision-event localisation of the C oracle against SciPy.

The exact mode (or_event_root: Illinois root of the gap on the step's cubic
Hermite interpolant) is pinned to the event root that scipy.integrate.solve_ivp
finds with DOP853 at rtol = atol = 1e-12 and a terminal event on the NumPy
restatement of the model (oracle/ckm1d_numpy.py) -- the reference's CPU
semantics (CKM1D_reference.py:702-714, 848-856) -- for the first collisions of
the 256-system plain workload, at the GPU threshold 0.1 and at the CPU
reference's 0.25 (CKM1D_reference.py:39).  The bracket mode (the reference
GPU / Numba rule) stops before the root with 0 < g < etol.
"""
import numpy as np
import pytest

from coupledbubble_control import workloads
from oracle import ckm1d_numpy as NP
from oracle import ckm1d_oracle as O


def _first_collisions(threshold, k):
    w = workloads.plain_cpu(seed=0)
    m = workloads.build_model(w)
    b = O.host_buffers_dict(m)
    sp = O.spec_from_model(m)
    sp.col_threshold = threshold
    O.solve(sp, b, kernel_steps=1 << 20, max_steps=1 << 24)
    ev = np.flatnonzero(b["status_flags"] == 2)
    assert ev.size >= k
    return w, ev[np.argsort(b["actual_time"][ev])][:k]


def _oracle_event(w, s, threshold, mode, etol, tol):
    m = workloads.build_model(w.subset(np.array([s])))
    b = O.host_buffers_dict(m)
    sp = O.spec_from_model(m)
    sp.col_threshold, sp.event_mode, sp.etol, sp.rtol, sp.atol = threshold, mode, etol, tol, tol
    O.solve(sp, b, kernel_steps=1 << 20, max_steps=1 << 24)
    return int(b["status_flags"][0]), float(b["actual_time"][0]), float(b["actual_event"][0])


def _scipy_event(w, s, threshold):
    m = workloads.build_model(w.subset(np.array([s])))
    up, gp, dp, mx, y0 = NP.from_reference_layout(O.host_buffers_dict(m), 0, 2, 2)
    sol = NP.integrate(y0, 5.0, up, gp, dp, mx, kind="SW_N", rtol=1e-12, atol=1e-12, threshold=threshold)
    assert len(sol.t_events[0]) == 1
    return float(sol.t_events[0][0])


@pytest.mark.parametrize("threshold,k", [(0.1, 3), (0.25, 2)])
def test_exact_event_matches_scipy_root(threshold, k):
    w, systems = _first_collisions(threshold, k)
    for s in systems:
        t_ref = _scipy_event(w, s, threshold)
        for tol, bar in ((1e-9, 1e-8), (1e-11, 1e-10)):
            flag, t, g = _oracle_event(w, s, threshold, "exact", 1e-12, tol)
            assert flag == 2 and abs(g) < 1e-12
            assert abs(t - t_ref) <= bar * t_ref, (s, tol, t, t_ref)


def test_bracket_mode_stops_before_the_root():
    w, systems = _first_collisions(0.1, 3)
    for s in systems:
        t_ref = _scipy_event(w, s, 0.1)
        flag, t, g = _oracle_event(w, s, 0.1, "bracket", 1e-5, 1e-9)
        assert flag == 2 and 0.0 < g < 1e-5
        assert t_ref * (1 - 1e-4) < t < t_ref


def test_infinite_radius_stops_as_non_finite():
    """An infinite radius makes the starting gap -inf, so the first trial step counts as a
    crossing and the bracket retry's factor inf/inf gives a NaN step size, which passes
    every clamp; it is flagged SINGULARITY_OR_NAN (-4) at the next crossing instead of
    being retried until the step budget runs out (65 536 attempts, ACTIVE, before the
    check).  A NaN radius stops the same way; the other systems are unaffected."""
    w = workloads.plain_cpu(seed=0).subset(np.arange(4))
    clean = O.host_buffers_dict(workloads.build_model(w))
    m = workloads.build_model(w)
    b = O.host_buffers_dict(m)
    b["actual_state"][0, 1, 0] = np.inf
    b["actual_state"][0, 2, 1] = np.nan
    sp = O.spec_from_model(m)
    r = O.solve(sp, b, kernel_steps=2048, max_steps=1 << 16)
    r0 = O.solve(O.spec_from_model(m), clean, kernel_steps=2048, max_steps=1 << 16)
    assert b["status_flags"][1] == -4 and b["status_flags"][2] == -4
    assert r["total_steps"][1] <= 4 and r["total_steps"][2] <= 4
    for s in (0, 3):
        assert b["status_flags"][s] == clean["status_flags"][s] >= 1
        assert r["total_steps"][s] == r0["total_steps"][s]
        np.testing.assert_array_equal(b["actual_state"][:, s], clean["actual_state"][:, s])
