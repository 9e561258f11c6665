"""Serving mode on the GPU: the persistent ring-fed kernel against the plain kernel
and the oracle, alone-vs-mix equality of the serving loop, relaunch after the
kernel idles out, and admission / validation errors.

A serving item runs the same attempt code as a plain launch from the same
buffers (t = 0, event value reset, amplitudes from the item), so serve and plain
agree bit for bit; the plain path is pinned to the oracle elsewhere and the
first test repeats that check for serving directly.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SERVE_TIMEOUT_S = 60.0


def _models(n, ns, seed, variant=None):
    from coupledbubble_control import workloads

    w = workloads.serving_bins(seed=seed, sizes={n: ns})[n]
    return w, workloads.build_model(w, variant=variant), workloads.build_model(w, variant=variant)


def _wait(q, ids, seq):
    t0 = time.time()
    while True:
        q.poll()
        if q.done(ids, seq).all():
            return
        if time.time() - t0 > SERVE_TIMEOUT_S:
            raise TimeoutError(f"serving items not finished after {SERVE_TIMEOUT_S} s")
        time.sleep(2e-4)


def _plain_step(m, ids, pa_pa, periods):
    """One plain launch of a step for systems ``ids`` (others parked)."""
    import torch

    from coupledbubble_control.backends.hip import rkck45_coupled_hip_solver as hs

    m._ensure_buffers()
    if m.runtime.solver is None:
        m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
    db = m.runtime.device_buffers
    dev = db.state.actual_state.device
    ti = torch.from_numpy(np.asarray(ids, dtype=np.int64)).to(dev)
    db.state.time_end.copy_(db.state.actual_time)
    for k in range(m._model_spec.nk):
        db.params.control_params[k].index_copy_(0, ti, torch.from_numpy(np.ascontiguousarray(pa_pa[k])).to(dev))
    db.state.actual_time.index_fill_(0, ti, 0.0)
    db.state.time_end.index_fill_(0, ti, float(periods))
    db.status.actual_event.index_fill_(0, ti, -1.0)
    db.status.status_flags.index_fill_(0, ti, 0)
    hs.launch(m.runtime.kernel, db, 1 << 20)
    torch.cuda.synchronize()


def _rows(m, ids):
    db = m.runtime.device_buffers
    n = m._execution_spec.ups
    ids = np.asarray(ids)
    return (db.state.actual_state.cpu().numpy()[:, ids, :n], db.state.time_step.cpu().numpy()[ids],
            db.state.actual_time.cpu().numpy()[ids], db.status.status_flags.cpu().numpy()[ids])


@pytest.mark.parametrize("n,ns,variant,exact", [(2, 256, None, False), (4, 128, None, False),
                                               (2, 128, "group", False), (4, 128, None, True),
                                               (8, 64, None, True), (12, 24, None, False)])
def test_serve_equals_plain_bitwise(n, ns, variant, exact):
    """A served step equals a plain launch of the same step bit for bit; also in the exact
    event mode (N = 8: the Gauss--Jordan kernel's exact build on the serving path) and
    with two lanes per bubble (N = 12)."""
    import dataclasses

    from coupledbubble_control.backends.hip.serve import ServeOptions, ServeQueue

    print(f"\n[serve==plain] N={n} ns={ns} variant={variant} exact={exact}", flush=True)
    _, ms, mp = _models(n, ns, seed=11, variant=variant)
    if exact:
        for m in (ms, mp):
            m._solver_spec = dataclasses.replace(m._solver_spec, event_mode="exact", etol=1e-12)
    rng = np.random.default_rng(5)
    with ServeQueue(ms, ServeOptions(capacity=1024, max_attempts=200000, idle_us=2000)) as q:
        steps = [np.arange(ns), rng.choice(ns, size=ns // 3, replace=False)]
        for k, ids in enumerate(steps):
            pa = rng.uniform(0.0, 1.5, size=(2, ids.size)) * 1e5
            seq = q.submit(ids, pa, 5.0)
            assert seq is not None
            _wait(q, ids, seq)
            _plain_step(mp, ids, pa, 5.0)
            ys, hs_, ts, fs = _rows(ms, ids)
            yp, hp, tp, fp = _rows(mp, ids)
            np.testing.assert_array_equal(fs, fp)
            np.testing.assert_array_equal(ts, tp)
            np.testing.assert_array_equal(hs_, hp)
            np.testing.assert_array_equal(ys, yp)
            st, misc = q.results(ids)
            np.testing.assert_array_equal(st, ys)
            np.testing.assert_array_equal(misc[:, 2].astype(np.int32), fs)
            np.testing.assert_array_equal(misc[:, 0], ts)
            print(f"  step {k}: {ids.size} systems, flags {np.unique(fs, return_counts=True)}", flush=True)


@pytest.mark.parametrize("n,ns", [(2, 128), (4, 64)])
def test_serve_reset_item_equals_plain_reset(n, ns):
    """A new episode staged in host memory and started by a reset item gives the device
    rows and end states of reset_systems followed by a plain launch."""
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip.serve import ServeOptions, ServeQueue
    from coupledbubble_control.serving import c0_scene

    print(f"\n[serve reset] N={n} ns={ns}", flush=True)
    _, ms, mp = _models(n, ns, seed=16)
    lam = workloads.wavelength(workloads.water_25c(), 25.0)
    rng = np.random.default_rng(8)
    ids = np.arange(ns)
    with ServeQueue(ms, ServeOptions(capacity=256, max_attempts=200000, idle_us=2000)) as q:
        pa = rng.uniform(0.0, 1.5, size=(2, ns)) * 1e5
        _wait(q, ids, q.submit(ids, pa, 2.0))
        _plain_step(mp, ids, pa, 2.0)
        rid = np.sort(rng.choice(ns, size=ns // 4, replace=False))
        R0, X0 = c0_scene(n, rid.size, np.random.default_rng(1), lam)
        q.stage(rid, *ms.scene_rows(rid, R0, X0))
        mp.reset_systems(rid, R0, X0, h0=1e-6)
        torch.cuda.synchronize()
        reset = np.isin(ids, rid)
        pa2 = rng.uniform(0.0, 1.5, size=(2, ns)) * 1e5
        _wait(q, ids, q.submit(ids, pa2, 2.0, np.where(reset, 1e-6, 0.0), reset))
        _plain_step(mp, ids, pa2, 2.0)
    for a, b in zip(_rows(ms, ids), _rows(mp, ids)):
        np.testing.assert_array_equal(a, b)
    for name in ("unit_params", "coupling_matrices"):
        np.testing.assert_array_equal(getattr(ms.runtime.device_buffers.params, name).cpu().numpy(),
                                      getattr(mp.runtime.device_buffers.params, name).cpu().numpy())


def test_serve_vs_oracle():
    """One served step of 5 periods against the C oracle from the same state."""
    import test_gpu_plain as P

    from coupledbubble_control.backends.hip.serve import ServeOptions, ServeQueue
    from oracle import ckm1d_oracle as O

    print("\n[serve vs oracle] N=2 ns=128", flush=True)
    _, m, _ = _models(2, 128, seed=12)
    ref = O.host_buffers_dict(m)
    spec = O.spec_from_model(m)
    rng = np.random.default_rng(6)
    ids = np.arange(128)
    pa = rng.uniform(0.0, 1.5, size=(2, 128)) * 1e5
    ref["control_params"][0:2] = pa
    ref["actual_time"][:] = 0.0
    ref["time_end"][:] = 5.0
    ref["actual_event"][:] = -1.0
    O.solve(spec, ref, kernel_steps=1 << 20, max_steps=1 << 20)
    with ServeQueue(m, ServeOptions(capacity=256, max_attempts=200000, idle_us=2000)) as q:
        seq = q.submit(ids, pa, 5.0)
        _wait(q, ids, seq)
    m.buffer_ops.sync_to_host(m.runtime.device_buffers.state, m.runtime.host_buffers.state)
    m.buffer_ops.sync_to_host(m.runtime.device_buffers.status, m.runtime.host_buffers.status)
    assert np.array_equal(m.runtime.host_buffers.status.status_flags, ref["status_flags"])
    err = P._assert_states(m, ref, 2)
    print(f"  max rel err {err:.2e}", flush=True)


def _engine_run(bins_w, reqs, mode, order="env"):
    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip.serve import ServeOptions
    from coupledbubble_control.serving import ServingEngine

    bins = {n: workloads.build_model(w) for n, w in bins_w.items()}
    eng = ServingEngine(bins, periods=5.0, mode=mode, order=order,
                        serve_opts=ServeOptions(max_attempts=200000, idle_us=5000))
    try:
        stats = eng.run(reqs, timeout_s=120.0)
    finally:
        eng.close()
    out = {}
    for r in eng.completed:
        y, h, t, f = _rows(bins[r.n], r.env_ids)
        out[r.rid] = (y, h, t, f, r.flags.copy(), np.array([r.periods]))
    return stats, out


def _fresh_requests(venvs, seed):
    from coupledbubble_control.serving import Request

    rng = np.random.default_rng(seed)
    return [Request(i, n, i, ids, rng.uniform(0.0, 1.5, size=(2, ids.size)), 0.0) for i, (n, ids) in enumerate(venvs)]


def test_serving_alone_vs_mix():
    """Every request's results are the same served alone, in the mix, and by the plain
    one-launch-per-batch path."""
    from coupledbubble_control import workloads
    from coupledbubble_control.serving import vector_envs

    print("\n[alone vs mix] bins N=2 x 384, N=4 x 192", flush=True)
    bins_w = workloads.serving_bins(seed=13, sizes={2: 384, 4: 192})
    venvs = vector_envs({2: 384, 4: 192}, np.random.default_rng(3), lo=64, hi=160)
    reqs = _fresh_requests(venvs, 21)
    st_mix, mix = _engine_run(bins_w, reqs, "continuous")
    print(f"  mix: {st_mix}", flush=True)
    assert st_mix["requests"] == len(reqs)
    _, plain = _engine_run(bins_w, _fresh_requests(venvs, 21), "plain")
    for rid in mix:
        for a, b in zip(mix[rid], plain[rid]):
            np.testing.assert_array_equal(a, b)
    for pick in (0, len(reqs) - 1):
        one = [r for r in _fresh_requests(venvs, 21) if r.rid == pick]
        _, alone = _engine_run(bins_w, one, "continuous")
        for a, b in zip(mix[pick], alone[pick]):
            np.testing.assert_array_equal(a, b)
        print(f"  request {pick} ({reqs[pick].size} envs, N={reqs[pick].n}) alone == mix", flush=True)


def test_env_order_equals_vector_env_order():
    """Three steps of every vector env queued at once.  Ordered per env (an env's next step
    starts when its own previous one has finished, so a vector env's steps overlap) and
    ordered per vector env (a step starts when the whole previous one has finished), every
    request's flags and integrated periods and the final resident rows are the same bit for
    bit, and equal to three plain static batches."""
    from coupledbubble_control import workloads
    from coupledbubble_control.serving import Request, vector_envs

    print("\n[env order vs vector-env order] bins N=2 x 384, N=4 x 192, 3 steps per vector env", flush=True)
    sizes = {2: 384, 4: 192}
    bins_w = workloads.serving_bins(seed=17, sizes=sizes)
    venvs = vector_envs(sizes, np.random.default_rng(5), lo=64, hi=160)

    def requests():
        rng = np.random.default_rng(23)
        return [Request(k * len(venvs) + v, n, v, ids, rng.uniform(0.0, 1.5, size=(2, ids.size)), 0.0)
                for k in range(3) for v, (n, ids) in enumerate(venvs)]

    st_env, by_env = _engine_run(bins_w, requests(), "continuous", order="env")
    st_venv, by_venv = _engine_run(bins_w, requests(), "continuous", order="venv")
    _, by_plain = _engine_run(bins_w, requests(), "plain")
    print(f"  env order: {st_env['wall_s']:.3f} s, vector-env order: {st_venv['wall_s']:.3f} s", flush=True)
    assert st_env["requests"] == st_venv["requests"] == 3 * len(venvs)
    for rid in by_env:
        for a, b, c in zip(by_env[rid], by_venv[rid], by_plain[rid]):
            np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(a, c)


def test_serve_relaunch_after_idle_and_admission():
    from coupledbubble_control.backends.hip import _abi
    from coupledbubble_control.backends.hip.serve import ServeOptions, ServeQueue

    print("\n[relaunch/admission] N=2 ns=256", flush=True)
    _, m, _ = _models(2, 256, seed=14)
    for bad in (0, (1 << 20) + 1):   # the kernels shift the interval left by up to 3
        with pytest.raises(_abi.CBCError):
            ServeQueue(m, ServeOptions(capacity=128, poll_interval=bad))
    with ServeQueue(m, ServeOptions(capacity=128, max_attempts=200000, idle_us=500)) as q:
        ids = np.arange(64)
        pa = np.full((2, 64), 0.5e5)
        seq = q.submit(ids, pa, 5.0)   # five periods: thousands of attempts, so still running below
        # a system with an unfinished item is refused
        with pytest.raises(_abi.CBCError):
            q.submit(np.arange(0, 1), np.full((2, 1), 0.5e5), 1.0)
        # ring full: 64 of 128 slots are in use until these items finish
        assert q.submit(np.arange(64, 129), np.full((2, 65), 0.5e5), 1.0) is None
        _wait(q, ids, seq)
        time.sleep(0.05)   # > idle_us: the kernel idles out
        q.poll()
        assert q.stats()[0] == 1
        with pytest.raises(_abi.CBCError):
            q.submit(np.array([70, 70]), np.full((2, 2), 0.5e5), 1.0)   # duplicate system
        ids2 = np.arange(64, 128)
        seq2 = q.submit(ids2, pa, 1.0)
        assert seq2 is not None
        _wait(q, ids2, seq2)
        launches, published, _ = q.stats()
        assert launches == 2 and published == 128
        flags = q.results(ids2)[1][:, 2]
        assert set(np.unique(flags)).issubset({1.0, 2.0, -2.0})  # SUCCESS, collision, h below resolution


def test_serving_loop_with_resets_runs():
    """A short Poisson run of the serving loop with episode resets completes every request."""
    from coupledbubble_control import workloads
    from coupledbubble_control.backends.hip.serve import ServeOptions
    from coupledbubble_control.serving import ServingEngine, c0_scene, make_requests, vector_envs

    print("\n[serving loop] N=2 x 1024, N=4 x 512, 0.5 s of arrivals", flush=True)
    sizes = {2: 1024, 4: 512}
    bins_w = workloads.serving_bins(seed=15, sizes=sizes)
    venvs = vector_envs(sizes, np.random.default_rng(4))
    reqs = make_requests(venvs, rate_env_steps=20000.0, duration=0.5, seed=7)
    lam = workloads.wavelength(bins_w[2].mat, 25.0)
    bins = {n: workloads.build_model(w) for n, w in bins_w.items()}
    eng = ServingEngine(bins, periods=5.0, mode="continuous", reset_prob=0.05,
                        reset_scene=lambda n, k, rng: c0_scene(n, k, rng, lam),
                        serve_opts=ServeOptions(max_attempts=200000, idle_us=5000))
    try:
        st = eng.run(reqs, timeout_s=120.0)
    finally:
        eng.close()
    print(f"  {st}", flush=True)
    assert st["requests"] == len(reqs)
    assert st["resets"] > 0
    for m in bins.values():
        y = m.runtime.device_buffers.state.actual_state.cpu().numpy()
        assert np.isfinite(y).all()


def test_served_vec_env_equals_blocking_env_step():
    """Three env steps through ServedVecEnv give the observations and failures of the
    reference's blocking step (set_action, model.solve, observe, check_termination,
    advance_time, reset_envs) on the plain solver, bit for bit."""
    import torch

    from coupledbubble_control import workloads
    from coupledbubble_control.envs import ServedVecEnv
    from coupledbubble_control.workloads import sample_positions

    print("\n[served env] N=2 ns=256, 3 steps with resets", flush=True)
    w, ms, mp = _models(2, 256, seed=17)
    lam = workloads.wavelength(w.mat, 25.0)
    gap_um = 3.0 * float(w.R0.max())

    def scene(k, rng):
        return sample_positions(rng, k, 2, 0.0, 0.25, np.full(k, gap_um * 1e-6 / lam))

    env = ServedVecEnv(ms, time_step_length=2.0, scene=scene, seed=4)
    rng_ref = np.random.default_rng(4)
    rng_a = np.random.default_rng(9)
    mp._ensure_buffers()
    db = mp.runtime.device_buffers
    db.state.time_end.fill_(2.0)
    total_fail = 0
    try:
        for step in range(3):
            pa = rng_a.uniform(0.0, 1.5, size=(256, 2))
            obs, fail = env.step(pa)
            # the reference step on the plain path
            db.params.control_params[0:2].copy_(torch.from_numpy(np.ascontiguousarray(pa.T * 1.0e5)).cuda())
            mp.solve(max_steps=1 << 20, kernel_steps=1 << 20)
            ref_obs = db.state.actual_state[1, :, :2].to(torch.float32).cpu().numpy()
            ref_fail = (db.status.status_flags != 1).cpu().numpy()
            db.status.status_flags.zero_()
            db.state.actual_time.zero_()
            db.state.time_step.fill_(1e-6)
            db.status.actual_event.fill_(-1.0)
            ids = np.nonzero(ref_fail)[0]
            if ids.size:
                mp.reset_systems(ids, None, scene(ids.size, rng_ref), h0=1e-6)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(fail, ref_fail)
            np.testing.assert_array_equal(obs, ref_obs)
            total_fail += int(fail.sum())
            print(f"  step {step}: {int(fail.sum())} failures", flush=True)
    finally:
        env.close()
    assert total_fail > 0   # the reset path was exercised


def test_overlapped_rollout_equals_synchronous():
    """Four rollout steps of 256 envs (N = 2) under a deterministic policy: the envs split
    into three groups that step on their own, and every env stepping on its own (policy
    batches of >= 16 finished envs), see the observations, actions and failures of one
    group and of the synchronous ServedVecEnv.step loop, bit for bit."""
    from coupledbubble_control.envs import ServedVecEnv
    from coupledbubble_control.rollout import OverlappedRollout

    def policy(o, ids):
        x = o.astype(np.float64)
        return np.clip(0.75 + 0.6 * np.sin(40.0 * x[:, :2] + 0.01 * ids[:, None]), 0.0, 1.5)

    steps = 4
    got = {}
    for groups in (0, 1, 3, "env"):   # 0: the synchronous loop
        print(f"\n[rollout] N=2 ns=256, {steps} steps, groups {groups}", flush=True)
        w, m, _ = _models(2, 256, seed=31)
        obs0 = m.runtime.host_buffers.state.actual_state[1, :, :2].astype(np.float32)
        env = ServedVecEnv(m, time_step_length=2.0, scene=None)
        try:
            if groups == 0:
                ids = np.arange(256)
                obs, acts, fails = [obs0], [], []
                for _ in range(steps):
                    a = policy(obs[-1], ids)
                    o, f = env.step(a)
                    obs.append(o)
                    acts.append(a)
                    fails.append(f)
                got[groups] = {"observations": np.stack(obs), "actions": np.stack(acts), "failures": np.stack(fails)}
            else:
                got[groups] = OverlappedRollout(env, policy, groups=groups, min_batch=16).collect(obs0, steps)
        finally:
            env.close()
    for groups in (1, 3, "env"):
        for k in ("observations", "actions", "failures"):
            np.testing.assert_array_equal(got[groups][k], got[0][k], err_msg=f"groups {groups}: {k}")
