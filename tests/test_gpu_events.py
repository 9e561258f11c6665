"""Exact collision-event mode on the GPU against the oracle's exact mode.

Lane mapping (256 x N2 plain workload, 5 periods, 12 collisions) and group
mapping (the same systems, N=3 and N=8 systems of the mixed workload, and
N=16 systems drawn the same way; the Gauss--Jordan kernels for N = 8, 16 run
the exact mode in builds of their own): flags equal, event times within 1e-9
relative (both run the same retries; they
differ by the rounding of the RK states), every gap at the end within etol,
states of the systems that finished normally within the parity bar of
test_gpu_plain.py.  The oracle's exact mode
is pinned to SciPy's event root in tests/test_oracle_events.py.
"""
import numpy as np
import pytest

from test_gpu_plain import ATOL_FLOOR, FLOOR_FRAC, RTOL

pytestmark = pytest.mark.gpu


def _exact(model):
    import dataclasses

    model._solver_spec = dataclasses.replace(model._solver_spec, event_mode="exact", etol=1e-12)
    return model


def _pair(w, variant):
    from coupledbubble_control import workloads
    from oracle import ckm1d_oracle as O

    m = _exact(workloads.build_model(w, variant=variant))
    ref = O.host_buffers_dict(m)
    O.solve(O.spec_from_model(m), ref, kernel_steps=1 << 20, max_steps=1 << 24)
    m.solve(max_steps=1 << 24, kernel_steps=1 << 20, sync_to_host=True)
    return m, ref


def _check(m, ref, n):
    hb = m.runtime.host_buffers
    flags = hb.status.status_flags
    np.testing.assert_array_equal(flags, ref["status_flags"])
    ev = flags == 2
    assert ev.sum() > 0
    np.testing.assert_allclose(hb.state.actual_time[ev], ref["actual_time"][ev], rtol=1e-9)
    assert np.all(np.abs(hb.status.actual_event[ev]) < 1e-12)
    keep = flags >= 1       # failures (STEP_SIZE_LIMIT at a wall speed near c) are compared by flag only
    got, want = hb.state.actual_state[:, keep, :n], ref["actual_state"][:, keep, :n]
    scale = (np.maximum(np.abs(want).max(axis=2, keepdims=True),
                        FLOOR_FRAC * np.abs(want).max(axis=(1, 2), keepdims=True)) + ATOL_FLOOR)
    assert (np.abs(got - want) / scale).max() <= RTOL
    return int(ev.sum())


@pytest.mark.parametrize("variant", ["lane", "group"])
def test_exact_events_plain_workload(device, variant):
    from coupledbubble_control import workloads

    print(f"case: exact events plain 256xN2 {variant}", flush=True)
    m, ref = _pair(workloads.plain_cpu(seed=0), variant)
    print(f"  {_check(m, ref, 2)} events", flush=True)


def test_exact_events_three_bubbles(device):
    from coupledbubble_control import workloads

    print("case: exact events N=3 (mixed workload bin), 5 periods", flush=True)
    w3 = [w for w in workloads.mixed_uneven() if w.n == 3][0]
    w3 = w3.subset(np.arange(512))
    w3.t_end = np.minimum(w3.t_end, 5.0)
    m, ref = _pair(w3, "group")
    print(f"  {_check(m, ref, 3)} events", flush=True)


def _sixteen_bubbles(ns=64, seed=16):
    """N=16 systems with the mixed workload's draws: log-uniform R0 in [10, 60] um,
    sorted positions on [0, lambda/2] with gaps >= 3 max(R0), 25 + 50 kHz at
    PA ~ U[0, 1.5] bar, 5 periods (the oracle ends 35 of 64 at a collision)."""
    from coupledbubble_control.workloads import Workload, sample_positions, water_25c, wavelength

    rng = np.random.default_rng(seed)
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    n = 16
    R0 = np.exp(rng.uniform(np.log(10.0), np.log(60.0), size=(ns, n)))
    X0 = sample_positions(rng, ns, n, 0.0, 0.5, 3.0 * R0.max(axis=1) * 1e-6 / lam)
    PA = rng.uniform(0.0, 1.5, size=(2, ns))
    FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
    return Workload("exact_N16", n, R0, X0, PA, FR, np.zeros((2, ns)), np.full(ns, 5.0), mat=mat)


@pytest.mark.parametrize("n", [8, 16])
def test_exact_events_gauss_jordan(device, n):
    from coupledbubble_control import workloads

    print(f"case: exact events N={n} (Gauss-Jordan group kernel), 5 periods", flush=True)
    if n == 8:
        w = [w for w in workloads.mixed_uneven() if w.n == 8][0].subset(np.arange(256))
        w.t_end = np.minimum(w.t_end, 5.0)
    else:
        w = _sixteen_bubbles()
    m, ref = _pair(w, "group")
    info = m.runtime.kernel.info()
    assert info.mapping == "group"
    assert n == 16 or info.local_bytes == 0   # N = 16 spills in either build
    print(f"  {_check(m, ref, n)} events", flush=True)
