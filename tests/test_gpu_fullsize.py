"""The four GPU configurations at their full sizes, over short horizons.

At full size the oracle would take minutes, so parity is shown by properties
that do not depend on size: every launch form that only changes where and when
a system runs (persistent work queue, mixed-N ensemble, serving kernel) gives
the plain launch's results bit for bit on the whole ensemble, and a sample of
systems spread over the ensemble agrees with the oracle to the parity bar of
test_gpu_plain.py (final states 1e-6 relative, flags equal, attempt counts 2 %).
"""
import numpy as np
import pytest

from test_gpu_plain import ATOL_FLOOR, FLOOR_FRAC, RTOL, check_counts

pytestmark = pytest.mark.gpu


def _device_rows(m):
    d = m.runtime.device_buffers
    return {"state": d.state.actual_state.cpu().numpy(), "t": d.state.actual_time.cpu().numpy(),
            "h": d.state.time_step.cpu().numpy(), "flags": d.status.status_flags.cpu().numpy(),
            "steps": d.status.total_steps.cpu().numpy(), "rejected": d.status.rejected_steps.cpu().numpy()}


def _assert_equal_rows(a, b, what):
    for k in a:
        bad = np.flatnonzero((a[k] != b[k]).reshape(a[k].shape[0], -1).any(axis=1)) if a[k].ndim == 1 else \
            np.flatnonzero((a[k] != b[k]).any(axis=(0, 2)))
        if bad.size:
            rows = [f"{s}: flags {a['flags'][s]}/{b['flags'][s]} t {a['t'][s]!r}/{b['t'][s]!r} "
                    f"h {a['h'][s]!r}/{b['h'][s]!r} steps {a['steps'][s]}/{b['steps'][s]}" for s in bad[:6]]
            raise AssertionError(f"{what}: {k} differs for {bad.size} systems, e.g.\n  " + "\n  ".join(rows))


def _oracle_sample(w, idx, rows, n):
    """Oracle on the sampled systems vs the GPU rows of those systems."""
    from coupledbubble_control import workloads
    from oracle import ckm1d_oracle as O

    sub = w.subset(idx)
    m = workloads.build_model(sub)
    ref = O.host_buffers_dict(m)
    cnt = O.solve(O.spec_from_model(m), ref, kernel_steps=1 << 20, max_steps=1 << 24)
    np.testing.assert_array_equal(rows["flags"][idx], ref["status_flags"])
    keep = ref["status_flags"] >= 0
    got = rows["state"][:, idx][:, keep, :n]
    want = ref["actual_state"][:, keep, :n]
    scale = (np.maximum(np.abs(want).max(axis=2, keepdims=True), FLOOR_FRAC * np.abs(want).max(axis=(1, 2),
                                                                                             keepdims=True))
             + ATOL_FLOOR)
    err = np.abs(got - want) / scale
    assert err.max() <= RTOL, f"max rel err {err.max():.3e}"
    check_counts(rows["steps"][idx], rows["rejected"][idx], cnt, failed=ref["status_flags"] < 0)
    return float(err.max())


def test_param_study_full_grid(device):
    """65 536 systems of N=2 (256 x 256 PA grid), 2 periods: queued == plain, sample vs oracle."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip import rkck45_coupled_hip_solver as hs

    print("case: param study 65536 x N2, 2 periods", flush=True)
    w = workloads.param_study()
    w.t_end[:] = 2.0
    a = workloads.build_model(w)
    a.solve(max_steps=1 << 22, kernel_steps=1 << 20)
    b = workloads.build_model(w)
    b._ensure_buffers()
    b.runtime.solver = b._make_solver(b.runtime.backend, b.runtime.variant)
    hs.solve_queued(b.runtime.kernel, b.runtime.device_buffers, refill=1, kernel_steps=1 << 20, max_steps=1 << 22)
    ra, rb = _device_rows(a), _device_rows(b)
    _assert_equal_rows(ra, rb, "queued vs plain")
    idx = np.linspace(0, w.ns - 1, 96).astype(int)
    err = _oracle_sample(w, idx, ra, 2)
    print(f"  flags {np.unique(ra['flags'], return_counts=True)}, sample err {err:.2e}", flush=True)


def test_mixed_full_ensemble(device):
    """32 768 systems, N in {2, 3, 4, 8}, horizons capped at 1 period: the ensemble (one
    persistent launch per bin, concurrent) == one plain solve per bin; samples vs oracle."""
    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import MixedEnsemble

    print("case: mixed 32768, horizons <= 1 period", flush=True)
    ws = workloads.mixed_uneven()
    for w in ws:
        w.t_end = np.minimum(w.t_end, 1.0)
    ens_models = [workloads.build_model(w) for w in ws]
    MixedEnsemble(ens_models).solve()
    for w, me in zip(ws, ens_models):
        mp = workloads.build_model(w)
        mp.solve(max_steps=1 << 22, kernel_steps=1 << 20)
        re_, rp = _device_rows(me), _device_rows(mp)
        _assert_equal_rows(re_, rp, f"ensemble vs plain N={w.n}")
        idx = np.linspace(0, w.ns - 1, 24 if w.n < 8 else 12).astype(int)
        err = _oracle_sample(w, idx, rp, w.n)
        print(f"  N={w.n} {w.ns} systems: sample err {err:.2e}", flush=True)


def test_serving_full_resident_set(device):
    """16 384 resident systems (12 288 x N=2, 4 096 x N=4), one 5-period step of every
    vector env: continuous == plain static batches; sample vs oracle."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip.serve import ServeOptions
    from coupledbubble_control.serving import Request, ServingEngine, vector_envs

    print("case: serving 16384 resident, one step of 33 vector envs", flush=True)
    sizes = {2: 12288, 4: 4096}
    venvs = vector_envs(sizes, np.random.default_rng(3))

    def requests():
        rng = np.random.default_rng(5)
        return [Request(i, n, i, ids, rng.uniform(0.0, 1.5, size=(2, ids.size)), 0.0)
                for i, (n, ids) in enumerate(venvs)]

    results = {}
    for mode in ("continuous", "plain"):
        bins_w = workloads.serving_bins(seed=3, sizes=sizes)
        bins = {n: workloads.build_model(w) for n, w in bins_w.items()}
        eng = ServingEngine(bins, periods=5.0, mode=mode, serve_opts=ServeOptions(max_attempts=1 << 20))
        try:
            st = eng.run(requests(), timeout_s=120.0)
        finally:
            eng.close()
        print(f"  {mode}: {st['wall_s']:.2f} s, {st['requests']} requests", flush=True)
        results[mode] = {n: _device_rows(m) for n, m in bins.items()}
    for n in sizes:
        _assert_equal_rows(results["continuous"][n], results["plain"][n], f"continuous vs plain N={n}")
    # oracle on samples of the same first step
    bins_w = workloads.serving_bins(seed=3, sizes=sizes)
    reqs = requests()
    for n, w in bins_w.items():
        pa = np.zeros((2, w.ns))
        for r in reqs:
            if r.n == n:
                pa[:, r.env_ids] = r.actions
        w.PA = pa
        w.t_end[:] = 5.0
        idx = np.linspace(0, w.ns - 1, 32).astype(int)
        err = _oracle_sample(w, idx, results["continuous"][n], n)
        print(f"  N={n}: sample err {err:.2e}", flush=True)


def test_large_coupling_full(device):
    """4 096 systems of N=16 mixed with 4 096 of N=2, 0.25 periods: ensemble == plain per
    bin; N=16 sample vs oracle."""
    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import MixedEnsemble

    print("case: large coupling 4096 x N16 + 4096 x N2, 0.25 periods", flush=True)
    ws = workloads.large_coupling()
    for w in ws:
        w.t_end[:] = 0.25
    ens_models = [workloads.build_model(w) for w in ws]
    MixedEnsemble(ens_models).solve()
    for w, me in zip(ws, ens_models):
        mp = workloads.build_model(w)
        mp.solve(max_steps=1 << 22, kernel_steps=1 << 20)
        _assert_equal_rows(_device_rows(me), _device_rows(mp), f"ensemble vs plain N={w.n}")
    w16 = ws[0]
    idx = np.linspace(0, w16.ns - 1, 6).astype(int)
    err = _oracle_sample(w16, idx, _device_rows(ens_models[0]), 16)
    print(f"  N=16 sample err {err:.2e}", flush=True)
