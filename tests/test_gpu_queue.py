"""Persistent work-queue launches, mixed-N ensembles and graph replay.

A queued launch changes only which lane integrates which system and when, never
the arithmetic of a system, so its results must equal the plain launch bit for
bit whatever the queue order and refill threshold."""
import numpy as np
import pytest

from test_gpu_plain import _assert_counts, _assert_states

pytestmark = pytest.mark.gpu


def _uneven(n, ns, seed, hmax=3.0):
    from coupledbubble_control import workloads

    ws = {w.n: w for w in workloads.mixed_uneven(seed=seed, total=15 * 64)}
    w = ws[n]
    idx = np.arange(ns) % w.ns
    w = w.subset(idx)
    rng = np.random.default_rng(seed)
    w.t_end = rng.uniform(0.2, hmax, size=ns)
    return w


def _snapshot(m):
    d = m.runtime.device_buffers
    return {k: v.detach().cpu().numpy().copy() for k, v in (
        ("state", d.state.actual_state), ("t", d.state.actual_time), ("h", d.state.time_step),
        ("flags", d.status.status_flags), ("steps", d.status.total_steps), ("event", d.status.actual_event))}


@pytest.mark.parametrize("n,variant", [(2, "lane"), (2, "group"), (3, "group"), (8, "group")])
@pytest.mark.parametrize("refill", [1, 64])
def test_queued_equals_plain(device, n, variant, refill):
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant
    from coupledbubble_control.backends.hip import rkck45_coupled_hip_solver as hs

    ns = {2: 200, 3: 90, 8: 24}[n]
    print(f"case: queued vs plain N={n} {variant} refill={refill} ns={ns}", flush=True)
    w = _uneven(n, ns, seed=30 + n)
    a = workloads.build_model(w, variant=KernelVariant(variant))
    a.solve(max_steps=10 ** 7, kernel_steps=10 ** 6)
    b = workloads.build_model(w, variant=KernelVariant(variant))
    b._ensure_buffers()
    b.runtime.solver = b._make_solver(b.runtime.backend, b.runtime.variant)
    gen = torch.Generator().manual_seed(n)
    order = torch.randperm(ns, generator=gen).to(torch.int32).cuda()
    launches = hs.solve_queued(b.runtime.kernel, b.runtime.device_buffers, order, waves=3, refill=refill,
                               kernel_steps=10 ** 6, max_steps=10 ** 7)
    assert launches == 1
    sa, sb = _snapshot(a), _snapshot(b)
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_queued_budget_relaunch(device):
    """A small per-launch budget forces relaunches; the result still equals one plain solve."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip import rkck45_coupled_hip_solver as hs

    print("case: queued relaunch, budget 300", flush=True)
    w = _uneven(2, 128, seed=40)
    a = workloads.build_model(w)
    a.solve(max_steps=10 ** 7, kernel_steps=10 ** 6)
    b = workloads.build_model(w)
    b._ensure_buffers()
    b.runtime.solver = b._make_solver(b.runtime.backend, b.runtime.variant)
    launches = hs.solve_queued(b.runtime.kernel, b.runtime.device_buffers, kernel_steps=300, max_steps=10 ** 7)
    assert launches > 1
    sa, sb = _snapshot(a), _snapshot(b)
    for k in ("state", "t", "h", "flags", "steps"):
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_mixed_ensemble_against_oracle(device):
    """Bins N = 2, 3, 4, 8 with uneven horizons solved together; each bin vs the oracle."""
    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import MixedEnsemble
    from oracle import ckm1d_oracle as O

    print("case: mixed ensemble 2/3/4/8", flush=True)
    models, refs = [], []
    for n, ns in ((2, 160), (3, 80), (4, 40), (8, 20)):
        w = _uneven(n, ns, seed=50 + n, hmax=1.5)
        m = workloads.build_model(w)
        ref = O.host_buffers_dict(m)
        cnt = O.solve(O.spec_from_model(m), ref, kernel_steps=10 ** 6, max_steps=10 ** 7)
        models.append(m)
        refs.append((ref, cnt))
    ens = MixedEnsemble(models)
    rounds = ens.solve()
    assert rounds == 1
    for m, (ref, cnt) in zip(models, refs):
        m.buffer_ops.sync_to_host(m.runtime.device_buffers.state, m.runtime.host_buffers.state)
        m.buffer_ops.sync_to_host(m.runtime.device_buffers.status, m.runtime.host_buffers.status)
        np.testing.assert_array_equal(m.runtime.host_buffers.status.status_flags, ref["status_flags"])
        _assert_states(m, ref, m._execution_spec.ups)
        _assert_counts(m, cnt)


def test_graph_replay_matches_eager(device):
    """One launch round captured in a HIP graph and replayed three times (eager resets
    of the buffers in between) gives the eager result every time."""
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import MixedEnsemble

    print("case: graph capture and 3 replays", flush=True)

    def build():
        ms = []
        for n, ns in ((2, 96), (4, 32)):
            ms.append(workloads.build_model(_uneven(n, ns, seed=60 + n, hmax=1.0)))
        return ms

    eager = MixedEnsemble(build())
    eager.solve()
    want = [_snapshot(m) for m in eager.models]

    gm = build()
    ens = MixedEnsemble(gm)
    ens._prepare()
    init = [{k: v.clone() for k, v in (("state", m.runtime.device_buffers.state.actual_state),
                                         ("t", m.runtime.device_buffers.state.actual_time),
                                         ("h", m.runtime.device_buffers.state.time_step),
                                         ("event", m.runtime.device_buffers.status.actual_event))}
            for m in gm]
    ens.capture()
    for rep in range(3):
        for m, i0 in zip(gm, init):
            d = m.runtime.device_buffers
            d.state.actual_state.copy_(i0["state"])
            d.state.actual_time.copy_(i0["t"])
            d.state.time_step.copy_(i0["h"])
            d.status.actual_event.copy_(i0["event"])
            d.status.status_flags.zero_()
            d.status.total_steps.zero_()
        torch.cuda.synchronize()
        ens.replay()
        torch.cuda.synchronize()
        for m, wnt in zip(gm, want):
            got = _snapshot(m)
            for k in ("state", "t", "h", "flags", "steps"):
                np.testing.assert_array_equal(got[k], wnt[k], err_msg=f"replay {rep}: {k}")


def test_autotune_choices_and_tuned_ensemble(device):
    """The autotuner times every candidate (lane/group, refill, lane workgroup size) and
    the ensemble run with its choices equals one plain solve per bin bit for bit."""
    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import MixedEnsemble, autotune

    print("case: autotune N=2, N=4", flush=True)

    def build(variants=None):
        return [workloads.build_model(_uneven(n, ns, seed=70 + n, hmax=1.0),
                                      variant=KernelVariant(variants[n]) if variants else None)
                for n, ns in ((2, 256), (4, 64))]

    from coupledbubble_control.backends import KernelVariant

    tune = autotune(build(), horizon=0.2, limit=256, refills=(1, 64), blocks=(64, 128))
    assert set(tune) == {2, 4}
    assert len(tune[2].trials) == 2 * 2 + 2 and len(tune[4].trials) == 2
    assert tune[4].choice.variant == "group"
    assert all(t > 0 for r in tune.values() for t in r.trials.values())
    models = build()
    MixedEnsemble(models, {n: r.choice for n, r in tune.items()}).solve()
    # a plain solve per bin with the mapping the tuner chose (lane and group differ in rounding)
    for m, mp in zip(models, build({n: r.choice.variant for n, r in tune.items()})):
        mp.solve(max_steps=10 ** 7, kernel_steps=10 ** 6)
        sa, sb = _snapshot(m), _snapshot(mp)
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    print(f"  choices { {n: r.choice for n, r in tune.items()} }", flush=True)


def _fused_bins(spec, seed):
    """Workloads of the bins (n, ns): uneven horizons; N = 16 from the large-coupling set."""
    from coupledbubble_control import workloads

    out = []
    for n, ns in spec:
        if n == 16:
            w = workloads.large_coupling(seed=seed, ns16=ns, ns2=4)[0]
            w.t_end = np.random.default_rng(seed).uniform(0.05, 0.3, size=ns)
        else:
            w = _uneven(n, ns, seed=seed + n, hmax=1.2)
        out.append(w)
    return out


@pytest.mark.parametrize("spec", [((8, 24), (4, 40), (3, 80), (2, 160)), ((16, 12), (2, 96)), ((2, 100), (4, 40))],
                         ids=["8-4-3-2", "16-2", "2-4"])
@pytest.mark.parametrize("budget", [1 << 20, 200], ids=["one-round", "budget-200"])
def test_fused_equals_per_bin(device, spec, budget):
    """The bins in one fused persistent kernel give, bit for bit, the per-bin queued
    launches' results (state, t, h, flags, attempts, event), also when a small attempt
    budget forces relaunch rounds over the systems still ACTIVE."""
    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import BinTuning, MixedEnsemble

    print(f"case: fused vs per-bin {spec} budget={budget}", flush=True)
    ws = _fused_bins(spec, seed=80)
    tune = {n: BinTuning("group", refill=1 + (n % 3)) for n, _ in spec}
    a = [workloads.build_model(w) for w in ws]
    ea = MixedEnsemble(a, tune, fused="off")
    ra = ea.solve(kernel_steps=budget)
    b = [workloads.build_model(w) for w in ws]
    eb = MixedEnsemble(b, tune, fused="on")
    rb = eb.solve(kernel_steps=budget)
    assert eb.fused_active and not ea.fused_active
    assert ra == rb
    if budget < 1000:
        assert ra > 1
    for ma, mb in zip(a, b):
        sa, sb = _snapshot(ma), _snapshot(mb)
        assert (sa["flags"] != 0).all()
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"N={ma._execution_spec.ups}: {k}")


def test_fused_selection_and_errors(device):
    """auto fuses only group-mapped bins that a fused kernel covers; "on" maps every bin to
    groups or refuses the set; the C ABI refuses a lane-mapped plan and leaves it untouched."""
    import ctypes

    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.backends import KernelVariant
    from coupledbubble_control.backends.hip import _abi
    from coupledbubble_control.backends.hip import rkck45_coupled_hip_solver as hs
    from coupledbubble_control.ensemble import BinTuning, MixedEnsemble

    print("case: fused selection", flush=True)
    w2, w4 = _uneven(2, 64, seed=91, hmax=0.5), _uneven(4, 16, seed=92, hmax=0.5)
    ens = MixedEnsemble([workloads.build_model(w2), workloads.build_model(w4)],
                        {2: BinTuning("lane"), 4: BinTuning("group")})
    ens._prepare()
    assert not ens.fused_active
    ens = MixedEnsemble([workloads.build_model(w2), workloads.build_model(w4)],
                        {2: BinTuning("lane"), 4: BinTuning("group")}, fused="on")
    ens._prepare()
    assert ens.fused_active and ens.models[0].runtime.kernel.info().mapping == "group"
    w5 = _uneven(8, 8, seed=93, hmax=0.5)
    w5.n = 5
    w5.R0, w5.X0 = w5.R0[:, :5], w5.X0[:, :5]
    m5 = workloads.build_model(w5)
    ens = MixedEnsemble([workloads.build_model(w2), m5])
    ens._prepare()
    assert not ens.fused_active
    with pytest.raises(ValueError, match="no fused kernel"):
        MixedEnsemble([workloads.build_model(w2), m5], fused="on")._prepare()
    # C ABI: a lane-mapped plan is refused before anything is launched
    ml = workloads.build_model(w2, variant=KernelVariant.LANE)
    ml._ensure_buffers()
    ml.runtime.solver = ml._make_solver(ml.runtime.backend, ml.runtime.variant)
    before = _snapshot(ml)
    order = hs.horizon_order(ml.runtime.device_buffers)
    with pytest.raises(_abi.CBCError) as ei:
        hs.launch_fused([ml.runtime.kernel], [ml.runtime.device_buffers], [order], [1])
    assert ei.value.code == _abi.CBC_ERR_UNSUPPORTED
    torch.cuda.synchronize()
    after = _snapshot(ml)
    for k in before:
        np.testing.assert_array_equal(before[k], after[k], err_msg=k)


@pytest.mark.parametrize("case", ["nk1-sw_a", "const-exact"])
def test_fused_other_specs(device, case):
    """The fused kernels for one harmonic and the other field types, and the exact
    collision-event mode through a fused launch: bit for bit the per-bin launches."""
    import dataclasses

    from coupledbubble_control import workloads
    from coupledbubble_control.ensemble import BinTuning, MixedEnsemble

    print(f"case: fused {case}", flush=True)
    spec = ((8, 16), (2, 96)) if case == "nk1-sw_a" else ((4, 40), (3, 48), (2, 96))
    ws = _fused_bins(spec, seed=95)
    for w in ws:
        if case == "nk1-sw_a":
            w.PA, w.FR, w.PS = w.PA[:1], w.FR[:1], w.PS[:1]
            w.field = "SW_A"
        else:
            w.field = "CONST"

    def build():
        ms = [workloads.build_model(w) for w in ws]
        if case == "const-exact":
            for m in ms:
                m._solver_spec = dataclasses.replace(m._solver_spec, event_mode="exact", etol=1e-12)
        return ms

    tune = {n: BinTuning("group") for n, _ in spec}
    a, b = build(), build()
    MixedEnsemble(a, tune, fused="off").solve()
    eb = MixedEnsemble(b, tune, fused="on")
    eb.solve()
    assert eb.fused_active
    for ma, mb in zip(a, b):
        sa, sb = _snapshot(ma), _snapshot(mb)
        assert (sa["flags"] != 0).all()
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"N={ma._execution_spec.ups}: {k}")
