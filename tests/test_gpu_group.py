"""GPU parity of the group mapping (one lane per bubble, LDS-staged Gauss--Jordan
coupling solve) against the C oracle, for the bubble counts of the mixed and
large-coupling workloads.  Same bars as test_gpu_plain.py."""
import numpy as np
import pytest

from test_gpu_plain import RTOL, _assert_counts, _assert_states, _run_pair

pytestmark = pytest.mark.gpu


def _mixed(n, ns, periods, seed=2):
    from coupledbubble_control import workloads

    ws = {w.n: w for w in workloads.mixed_uneven(seed=seed, total=15 * 64)}
    w = ws[n].subset(np.arange(ns))
    w.t_end[:] = periods
    return w


def _any_n(n, ns, periods, seed=9):
    """Mixed-workload style systems of any bubble count: R0 log-uniform in [10, 100] um,
    sorted positions on [0, lambda/2] with gaps >= 3 max(R0), PA ~ U[0, 1.5] bar."""
    from coupledbubble_control import workloads

    rng = np.random.default_rng(seed + n)
    lam = workloads.wavelength(workloads.water_25c(), 25.0)
    R0 = np.exp(rng.uniform(np.log(10.0), np.log(100.0), size=(ns, n)))
    X0 = workloads.sample_positions(rng, ns, n, 0.0, 0.5, 3.0 * R0.max(axis=1) * 1e-6 / lam)
    PA = rng.uniform(0.0, 1.5, size=(2, ns))
    FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
    return workloads.Workload(f"any_N{n}", n, R0, X0, PA, FR, np.zeros((2, ns)), np.full(ns, periods))


@pytest.mark.parametrize("n,ns,periods", [(2, 128, 2.0), (3, 96, 2.0), (4, 64, 1.5), (8, 32, 1.0)])
def test_group_mapping_parity(device, n, ns, periods):
    from coupledbubble_control.backends import KernelVariant

    print(f"case: group mapping N={n} {ns} systems {periods} periods", flush=True)
    w = _mixed(n, ns, periods)
    m, ref, cnt = _run_pair(w, variant=KernelVariant.GROUP)
    info = m.runtime.kernel.info()
    assert info.mapping == "group" and info.local_bytes == 0
    assert np.array_equal(m.runtime.host_buffers.status.status_flags, ref["status_flags"])
    _assert_states(m, ref, n)
    _assert_counts(m, cnt)


@pytest.mark.parametrize("n", [8, 16])
def test_non_finite_system_is_flagged_alone(device, n):
    """A NaN radius in one system and an infinite one in another (coupling entries that
    are infinite give the pivot search keys that are signalling NaNs until quieted; the
    starting gap -inf makes the first trial step a crossing whose retry factor is NaN)
    stop those systems with SINGULARITY_OR_NAN within a few attempts; every other system
    of their waves ends bit for bit as in a run without them (the Gauss--Jordan kernels
    reduce within each group).  max_steps bounds a regression to seconds."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant

    print(f"case: non-finite injection N={n} 8 systems", flush=True)
    w = _any_n(n, 8, 0.25, seed=21)
    runs = []
    for inject in (False, True):
        m = workloads.build_model(w, variant=KernelVariant.GROUP)
        if inject:
            m.runtime.host_buffers.state.actual_state[0, 1, 2] = np.nan
            m.runtime.host_buffers.state.actual_state[0, 6, 0] = np.inf
        m.solve(max_steps=200000, kernel_steps=2048, sync_to_host=True)
        hb = m.runtime.host_buffers
        assert m.runtime.kernel.info().mapping == "group"
        runs.append((hb.status.status_flags.copy(), hb.state.actual_state.copy(), hb.state.actual_time.copy()))
    (f0, y0, t0), (f1, y1, t1) = runs
    assert f1[1] == -4 and f1[6] == -4
    rest = np.array([s for s in range(8) if s not in (1, 6)])
    assert np.all(f0 != -4)
    np.testing.assert_array_equal(f1[rest], f0[rest])
    np.testing.assert_array_equal(y1[:, rest], y0[:, rest])
    np.testing.assert_array_equal(t1[rest], t0[rest])


@pytest.mark.parametrize("n,ns,periods", [(5, 48, 1.0), (6, 40, 1.0), (7, 32, 1.0), (9, 24, 0.5), (10, 16, 0.5),
                                          (11, 16, 0.5), (12, 16, 0.5), (13, 12, 0.5), (14, 12, 0.5), (15, 12, 0.5)])
def test_group_mapping_other_bubble_counts(device, n, ns, periods):
    """Bubble counts that are not powers of two run in groups with padding lanes in the
    Gauss--Jordan solve (N = 5..7 in 8 lanes; 9..15 in 32, two lanes per bubble)."""
    from coupledbubble_control.backends import KernelVariant

    print(f"case: group mapping N={n} {ns} systems {periods} periods", flush=True)
    w = _any_n(n, ns, periods)
    m, ref, cnt = _run_pair(w, variant=KernelVariant.GROUP)
    info = m.runtime.kernel.info()
    assert info.mapping == "group" and info.local_bytes == 0
    assert np.array_equal(m.runtime.host_buffers.status.status_flags, ref["status_flags"])
    _assert_states(m, ref, n)
    _assert_counts(m, cnt)


def test_group_n16_large_coupling(device):
    """N = 16 (32 x 32 coupling solve per RHS evaluation)."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant

    print("case: group mapping N=16, 12 systems, 0.25 periods", flush=True)
    w16 = workloads.large_coupling(seed=4, ns16=12, ns2=4)[0]
    w16.t_end[:] = 0.25
    m, ref, cnt = _run_pair(w16, variant=KernelVariant.GROUP)
    assert np.array_equal(m.runtime.host_buffers.status.status_flags, ref["status_flags"])
    _assert_states(m, ref, 16)
    _assert_counts(m, cnt)


def test_group_and_lane_agree_for_n2(device):
    """Both mappings of the same N=2 ensemble stay within the bar of each other."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant

    print("case: lane vs group N=2, 128 systems, 2 periods", flush=True)
    w = workloads.plain_cpu(seed=21, ns=128)
    w.t_end[:] = 2.0
    ml = workloads.build_model(w, variant=KernelVariant.LANE)
    mg = workloads.build_model(w, variant=KernelVariant.GROUP)
    for m in (ml, mg):
        m.solve(max_steps=10 ** 7, kernel_steps=4096, sync_to_host=True)
    a, b = ml.runtime.host_buffers, mg.runtime.host_buffers
    np.testing.assert_array_equal(a.status.status_flags, b.status.status_flags)
    want = a.state.actual_state[:, :, :2]
    scale = np.maximum(np.abs(want).max(axis=2, keepdims=True), 1e-3 * np.abs(want).max(axis=(1, 2), keepdims=True))
    assert np.all(np.abs(b.state.actual_state[:, :, :2] - want) <= RTOL * (scale + 1e-9))


def test_group_dense_and_windows(device):
    """Dense output and period windows through the group kernel."""
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant
    from oracle import ckm1d_oracle as O

    print("case: group N=4 windows, 32 systems, 3 periods", flush=True)
    w = _mixed(4, 32, 3.0)
    m = workloads.build_model(w, variant=KernelVariant.GROUP)
    ref = O.host_buffers_dict(m)
    out = O.solve_periodic(O.spec_from_model(m), ref, 1.0, 1, 2)
    max_r, mean_x = m.solve_periodic(period=1.0, first_window=1, n_windows=2)
    torch.cuda.synchronize()
    flags = m.runtime.device_buffers.status.status_flags.cpu().numpy()
    np.testing.assert_array_equal(flags, ref["status_flags"])
    ok = flags == 1
    mr = max_r.cpu().numpy()[:, ok, :4]
    rel = np.abs(mr - out["max_r"][:, ok, :4]) / out["max_r"][:, ok, :4]
    assert np.mean(rel <= 1e-6) >= 0.95 and rel.max() <= 2e-4
    np.testing.assert_allclose(mean_x.cpu().numpy()[:, ok, :4], out["mean_x"][:, ok, :4], rtol=1e-6, atol=1e-12)

    print("case: group N=3 dense output ndo=8", flush=True)
    w3 = _mixed(3, 32, 1.0)
    m3, ref3, _ = _run_pair(w3, ndo=8, variant=KernelVariant.GROUP)
    np.testing.assert_array_equal(m3.runtime.host_buffers.dense.dense_index, ref3["dense_index"])


@pytest.mark.parametrize("n", [12, 16])
def test_split_groups_dense_and_windows(device, n):
    """Two lanes per bubble (N = 9..16): only the first lane of a bubble stores, so dense
    rows, window statistics and end states match the oracle as for the other mappings."""
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant
    from oracle import ckm1d_oracle as O

    print(f"case: split group N={n} windows, 12 systems, 0.6 periods", flush=True)
    w = _any_n(n, 12, 0.6)
    m = workloads.build_model(w, variant=KernelVariant.GROUP)
    ref = O.host_buffers_dict(m)
    out = O.solve_periodic(O.spec_from_model(m), ref, 0.2, 1, 2)
    max_r, mean_x = m.solve_periodic(period=0.2, first_window=1, n_windows=2)
    torch.cuda.synchronize()
    flags = m.runtime.device_buffers.status.status_flags.cpu().numpy()
    np.testing.assert_array_equal(flags, ref["status_flags"])
    ok = flags == 1
    mr = max_r.cpu().numpy()[:, ok, :n]
    rel = np.abs(mr - out["max_r"][:, ok, :n]) / out["max_r"][:, ok, :n]
    assert np.mean(rel <= 1e-6) >= 0.95 and rel.max() <= 2e-4
    np.testing.assert_allclose(mean_x.cpu().numpy()[:, ok, :n], out["mean_x"][:, ok, :n], rtol=1e-6, atol=1e-12)
    assert np.all(max_r.cpu().numpy()[:, :, n:] == 0.0), "padding columns must stay untouched"

    print(f"case: split group N={n} dense output ndo=5", flush=True)
    md, refd, cnt = _run_pair(_any_n(n, 12, 0.4), ndo=5, variant=KernelVariant.GROUP)
    np.testing.assert_array_equal(md.runtime.host_buffers.dense.dense_index, refd["dense_index"])
    np.testing.assert_array_equal(md.runtime.host_buffers.status.status_flags, refd["status_flags"])
    _assert_states(md, refd, n)
    _assert_counts(md, cnt)


def test_auto_mapping_for_two_bubbles(device):
    """AUTO runs N = 2 in lane pairs while their waves fit one per SIMD (up to 32 768
    systems on 256 CUs) and one system per lane beyond."""
    import torch

    from coupledbubble_control import workloads

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    limit = 4 * cus * 64 // 2
    for ns, want in ((4096, "group"), (limit, "group"), (limit + 32, "lane"), (65536, "lane")):
        print(f"case: AUTO mapping, N=2, {ns} systems", flush=True)
        w = workloads.plain_cpu(seed=3, ns=ns)
        w.t_end[:] = 0.01
        m = workloads.build_model(w)
        m.solve(max_steps=1 << 20, kernel_steps=1 << 20)
        assert m.runtime.kernel.info().mapping == want, (ns, cus)


def test_no_solver_kernel_uses_scratch(device):
    """Every kernel the AUTO mapping selects (lane: N = 1, 2; group: N = 2 .. 16;
    one or two harmonics; the three field types) keeps its state in registers and LDS: no
    private (scratch) bytes per lane, with the exact-event block compiled into every
    kernel.  The lane kernels for N = 3, 4 (chosen only on request) spill, as recorded."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant

    print("case: scratch bytes of every solver kernel", flush=True)
    seen = []
    for variant, ns_list in ((KernelVariant.LANE, (1, 2, 3, 4)), (KernelVariant.GROUP, tuple(range(2, 17)))):
        for n in ns_list:
            for nk in (1, 2):
                for field in ("CONST", "SW_N", "SW_A"):
                    ns = 4
                    w = workloads.Workload(f"scratch_N{n}", n, np.full((ns, n), 30.0),
                                           np.tile(np.arange(n) / (2.0 * n), (ns, 1)),
                                           np.full((nk, ns), 0.5), np.full((nk, ns), 25.0), np.zeros((nk, ns)),
                                           np.full(ns, 0.1), field=field)
                    m = workloads.build_model(w, variant=variant)
                    m._ensure_buffers()
                    m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
                    info = m.runtime.kernel.info()
                    assert info.mapping == variant.value
                    seen.append((variant.value, n, nk, field, info.num_regs, info.local_bytes))
                    if variant == KernelVariant.GROUP or n <= 2:
                        assert info.local_bytes == 0, seen[-1]
    print(f"  {len(seen)} kernels, registers {min(s[4] for s in seen)}..{max(s[4] for s in seen)}; "
          f"spilling: {[s for s in seen if s[5]]}", flush=True)
