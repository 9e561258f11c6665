"""The environment kernels (cbc_env_*) and PosNBC1D on the GPU against the oracle.

Kernel level: set_action, the post-step kernel (observe, check_termination,
get_rewards, advance_time, done list) and the reset kernel against
oracle/env_step.py on the same inputs -- bit for bit, since both evaluate the
same float32 / float64 operations in the same order -- at a ragged size (the
last workgroup partial), the serving size (16 384 envs) and the largest
supported size (2^20 envs, 4096 workgroups in the done-list scan).

Environment level: PosNBC1D episodes against the oracle environment whose
physics is the C oracle solver; observations and rewards agree within 1e-6
absolute (float32 values of states that agree to ~1e-12), flags, done lists and
episode lengths exactly.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32 = np.float32
ACTION_ROWS, ACTION_SCALES, PA_COLS = [0, 1, 4, 5], [1e5, 1e5, np.pi, np.pi], [0, 1]


def _model(ne, n):
    from coupledbubble_control.models import CKM1D

    rng = np.random.default_rng(ne + n)
    return CKM1D(num_systems=ne, systems_per_block=64, units_per_system=n,
                 R0=rng.uniform(20, 80, size=(ne, n)), PA=np.full((2, ne), 0.5), FR=np.tile([[25.0], [50.0]], ne),
                 PS=np.zeros((2, ne)), REF_FREQ=25.0, acoustic_field="SW_N")


def _parts(n, rich):
    from oracle import env_step as E

    lo, hi = [-0.5 - 0.25 * (i % 2) for i in range(n)], [0.5 + 0.125 * i for i in range(n)]
    parts = [E.Part("XT", lo, hi), E.Part("X", lo, hi, stack=2)]
    if rich:
        parts += [E.Part("R", [0.0] * n, [4.0] * n), E.Part("U", [-8.0] * n, [8.0] * n, stack=3),
                  E.Part("Q", [-1.0] * n, [1.0] * n, stack=2)]
    return parts


def _kernels(model, parts, **kw):
    from coupledbubble_control.backends.hip import _abi
    from coupledbubble_control.backends.hip.env_step import EnvKernels, ObsPart
    from oracle import env_step as E

    kinds = {"XT": (_abi.CBC_ENV_PART_TARGET, 0)}
    kinds.update({k: (_abi.CBC_ENV_PART_STATE, v) for k, v in E.STATE_ROWS.items()})
    ops = [ObsPart(p.name, *kinds.get(p.name, (_abi.CBC_ENV_PART_FIXED, 0)), p.stack) for p in parts]
    bias, rscale = E.box_scale(parts)
    return EnvKernels(model, ops, torch.from_numpy(bias), torch.from_numpy(rscale), action_rows=ACTION_ROWS,
                      action_scales=ACTION_SCALES, pa_cols=PA_COLS, **kw)


ENV_PARAMS = dict(position_tolerance=0.05, dnorm=0.2, reward_exp=0.5, weights=(1.0, 0.3),
                  positive_terminal_reward=10.0, negative_terminal_reward=-10.0, apply_termination=True,
                  episode_length=3)


def _random_inputs(rng, ne, n, tile, xt):
    state = np.zeros((4, ne, tile))
    state[0, :, :n] = rng.uniform(0.5, 2.0, size=(ne, n))
    near = rng.uniform(-0.06, 0.06, size=(ne, n)) * (rng.random((ne, 1)) < 0.5)
    state[1, :, :n] = np.where(near != 0, xt + near, rng.uniform(-0.5, 0.5, size=(ne, n)))
    state[2:, :, :n] = rng.normal(size=(2, ne, n))
    status = rng.choice(np.array([1, 1, 1, 1, 0, -2, 2], dtype=np.int32), size=ne)
    t, h, ev = rng.uniform(0, 5, ne), rng.uniform(1e-5, 1e-2, ne), rng.uniform(-1, 1, ne)
    act = rng.uniform(0, 1, size=(ne, 4)).astype(F32)
    return state, t, h, ev, status, act


def _upload(model, state, t, h, ev, status):
    db = model.runtime.device_buffers
    db.state.actual_state.copy_(torch.from_numpy(state))
    db.state.actual_time.copy_(torch.from_numpy(t))
    db.state.time_step.copy_(torch.from_numpy(h))
    db.status.actual_event.copy_(torch.from_numpy(ev))
    db.status.status_flags.copy_(torch.from_numpy(status))


def _eq(a, b, what):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    np.testing.assert_array_equal(a, np.asarray(b), err_msg=what)


def _scene(rng, n, k, sp, align):
    """A random-gap bubble scene and fixed targets; returns the C scene and the oracle's positions."""
    from coupledbubble_control.backends.hip import _abi
    from oracle import env_step as E

    u = torch.from_numpy(rng.random((k, n - 1))).cuda()
    ua = torch.from_numpy(rng.random((k, 1))).cuda()
    sc = _abi.EnvSceneC()
    sc.lo, sc.available, sc.min_distance = sp.lo, sp.hi - sp.lo, sp.min_distance
    sc.room_per_gap = sp.max_distance - sp.min_distance
    sc.extra_gap_limit = (sp.hi - sp.lo) - (n - 1) * sp.min_distance
    sc.bubbles.method = _abi.CBC_SCENE_RANDOM
    sc.bubbles.random_align = int(align)
    sc.bubbles.u_gaps, sc.bubbles.u_align = u.data_ptr(), ua.data_ptr()
    tgt = rng.uniform(-0.4, 0.4, size=n)
    sc.targets.method = _abi.CBC_SCENE_FIXED
    for i, v in enumerate(tgt):
        sc.targets.base[i] = v
    spa = E.SceneParams(**{**sp.__dict__, "alignment": "random" if align else "center"})
    pos = E.random_uniform_positions(spa, u.cpu().numpy(), ua.cpu().numpy())
    return sc, (u, ua), pos, np.tile(tgt, (k, 1))


def _post_and_reset_case(ne, n, rich, steps, seed):
    from oracle import env_step as E

    rng = np.random.default_rng(seed)
    model = _model(ne, n)
    parts = _parts(n, rich)
    K = _kernels(model, parts, **ENV_PARAMS)
    O = E.EnvOracle(ne, n, parts, action_rows=ACTION_ROWS, action_scales=ACTION_SCALES, pa_cols=PA_COLS,
                    **ENV_PARAMS)
    tile = model._execution_spec.tile
    xt = rng.uniform(-0.4, 0.4, size=(ne, n))
    O.buffers["XT"].update(xt)
    K.part_view("XT").copy_(torch.from_numpy(xt.astype(F32)))
    sp = E.SceneParams(num_bubbles=n, domain_min=-0.5, domain_max=0.5, min_distance=0.05, max_distance=0.9)
    db = model.runtime.device_buffers
    for k in range(steps):
        state, t, h, ev, status, act = _random_inputs(rng, ne, n, tile, xt)
        _upload(model, state, t, h, ev, status)
        cp = db.params.control_params.cpu().numpy()
        O.set_action(act, cp)
        K.set_action(torch.from_numpy(act).cuda())
        _eq(db.params.control_params, cp, "control_params")
        counts = K.post(k)
        r = O.post(state, t, h, ev, status, act)
        out = K.outputs
        _eq(out.observations, r["observations"], f"observations step {k}")
        _eq(out.rewards, r["rewards"], f"rewards step {k}")
        for name in ("failure", "success", "terminated", "time_outs"):
            _eq(getattr(out, name), r[name], f"{name} step {k}")
        done = r["done_ids"]
        assert (counts.done, counts.failures, counts.successes, counts.time_outs) == (
            done.size, int(r["failure"].sum()), int(r["success"].sum()), int(r["time_outs"].sum()))
        _eq(K.done_ids[:done.size], done, "done ids")
        _eq(db.state.actual_time, t, "actual_time")
        _eq(db.state.time_step, h, "time_step")
        _eq(db.status.actual_event, ev, "actual_event")
        _eq(db.status.status_flags, status, "status")
        _eq(K.total_rewards, O.total_rewards, "total_rewards")
        _eq(K.algo_steps, O.algo_steps, "algo_steps")
        for p in parts:
            want = O.buffers[p.name].buffer
            _eq(K.part_view(p.name).reshape(want.shape), want, f"history {p.name}")
        if done.size:
            sc, keep, pos, tgt = _scene(rng, n, done.size, sp, align=bool(k % 2))
            fo, fr, fl = K.reset(sc, K.done_ids[:done.size], k + 1, final=True)
            state_after = db.state.actual_state.cpu().numpy().copy()
            fin = O.reset_envs(done, pos, tgt, state)
            _eq(fo, fin["final_observation"], "final_observation")
            _eq(fr, fin["episode_return"], "episode_return")
            _eq(fl, fin["episode_length"], "episode_length")
            _eq(state_after, state, "state after reset")
            _eq(K.observations, fin["observations"], "observations after reset")
            _eq(K.algo_steps, O.algo_steps, "algo_steps after reset")
            for p in parts:
                want = O.buffers[p.name].buffer
                _eq(K.part_view(p.name).reshape(want.shape), want, f"history {p.name} after reset")
        xt = O.buffers["XT"].get().astype(np.float64)
    torch.cuda.synchronize()


@pytest.mark.parametrize("ne,n,rich", [(1000, 3, True), (1, 2, False), (300, 16, True)])
def test_post_and_reset_kernels_match_oracle(device, ne, n, rich):
    print(f"case: env kernels ne={ne} n={n} rich={rich}", flush=True)
    _post_and_reset_case(ne, n, rich, steps=4, seed=ne + 7 * n)


def test_post_and_reset_kernels_serving_size(device):
    print("case: env kernels ne=16384 n=2", flush=True)
    _post_and_reset_case(16384, 2, False, steps=3, seed=5)


def test_post_kernel_largest_size(device):
    """2^20 envs (4096 workgroups): the last-workgroup scan orders every done id."""
    print("case: env kernels ne=2^20 n=1", flush=True)
    _post_and_reset_case(1 << 20, 1, False, steps=2, seed=9)


def test_scene_positions_kernel_matches_oracle(device):
    from coupledbubble_control.backends.hip import _abi
    from coupledbubble_control.backends.hip.env_step import scene_positions
    from oracle import env_step as E

    rng = np.random.default_rng(1)
    for n in (1, 2, 4, 16):
        sp = E.SceneParams(num_bubbles=n, domain_min=-0.5, domain_max=0.5, min_distance=0.02, max_distance=0.3)
        for align in (False, True):
            sc, keep, pos, _ = _scene(rng, n, 777, sp, align) if n > 1 else (None, None, None, None)
            if n == 1:
                sc = _abi.EnvSceneC()
                sc.lo, sc.available, sc.min_distance = sp.lo, sp.hi - sp.lo, sp.min_distance
                sc.room_per_gap, sc.extra_gap_limit = sp.max_distance - sp.min_distance, sp.hi - sp.lo
                ua = torch.from_numpy(rng.random((777, 1))).cuda()
                keep = ua
                sc.bubbles.method, sc.bubbles.random_align = _abi.CBC_SCENE_RANDOM, int(align)
                sc.bubbles.u_align = ua.data_ptr()
                spa = E.SceneParams(**{**sp.__dict__, "alignment": "random" if align else "center"})
                pos = E.random_uniform_positions(spa, np.zeros((777, 0)), ua.cpu().numpy())
                sc.targets.method = _abi.CBC_SCENE_FIXED
            got, _ = scene_positions(sc, 777, n, "cuda")
            _eq(got, pos, f"random scene n={n} align={align}")
        if n < 2:
            continue
        # equidistant: no distance (linspace), a distance centred, a distance at a random start
        for distance, align in ((None, False), (0.01, False), (0.01, True)):
            sc = _abi.EnvSceneC()
            sc.lo = sp.lo
            sc.bubbles.method = _abi.CBC_SCENE_EQUIDISTANT
            spa = E.SceneParams(**{**sp.__dict__, "alignment": "random" if align else "center",
                                   "min_distance": 0.001, "max_distance": 1.0})
            ua = torch.from_numpy(rng.random((50, 1))).cuda()
            want = E.equidistant_positions(spa, distance, 50, ua.cpu().numpy())
            base = E.torch_linspace(sp.lo, sp.hi, n) if distance is None else np.arange(n) * distance
            for i, v in enumerate(base):
                sc.bubbles.base[i] = v
            if distance is not None:
                span = (n - 1) * distance
                sc.bubbles.offset = (sp.lo + sp.hi - span) * 0.5
                sc.bubbles.max_start = sp.hi - span
                sc.bubbles.random_align = int(align)
                sc.bubbles.u_align = ua.data_ptr()
            sc.targets.method = _abi.CBC_SCENE_FIXED
            got, _ = scene_positions(sc, 50, n, "cuda")
            _eq(got, want, f"equidistant n={n} distance={distance} align={align}")


# ---------------------------------------------------------------------------
# PosNBC1D against the oracle environment

def _oracle_env(env, num_envs=None):
    from oracle import env_step as E

    parts = [E.Part(c.name, c.min, c.max, c.stack) for c in env.observation_space_schema.components]
    rp = env.reward_params
    ne = env.num_envs if num_envs is None else num_envs
    return E.EnvOracle(ne, env.num_bubbles, parts, action_rows=env._action_map._dev_cols,
                       action_scales=env._action_map._scales, pa_cols=list(env._action_map._feat_col["PA"]),
                       position_tolerance=env.position_tolerance, dnorm=1.0 / rp["r_dnorm"],
                       reward_exp=rp["reward_exp"], weights=rp["weights"].cpu().tolist(),
                       positive_terminal_reward=rp["positive_terminal_reward"],
                       negative_terminal_reward=rp["negative_terminal_reward"],
                       apply_termination=env.apply_termination, episode_length=env.episode_length)


def _oracle_scene(env, draws):
    from oracle import env_step as E

    g = env.scene_generator
    sp = E.SceneParams(num_bubbles=g.num_bubbles, domain_min=g.domain_min, domain_max=g.domain_max,
                       margin=g.margin, min_distance=g.min_distance, max_distance=g.max_distance,
                       alignment=g.alignment)
    d = {k: v.cpu().numpy() for k, v in draws.items()}
    pos = E.random_uniform_positions(sp, d["bubbles_gaps"], d.get("bubbles_align"))
    tgt = E.random_uniform_positions(sp, d["targets_gaps"], d.get("targets_align"))
    return pos, tgt


@pytest.mark.parametrize("n,alignment", [(2, "center"), (3, "random")])
def test_env_episodes_match_oracle(device, n, alignment):
    from coupledbubble_control.envs import PosNBC1D
    from oracle import ckm1d_oracle as C

    print(f"case: PosNBC1D n={n} alignment={alignment}, 64 envs, 5 steps", flush=True)
    ne = 64
    env = PosNBC1D(num_bubbles=n, num_envs=ne, envs_per_block=64, R0=[40.0, 60.0, 50.0][:n], PA=[1.0, 0.5],
                   time_step_length=2.0, episode_length=3, position_tolerance=0.05, alignment=alignment,
                   seed=17, verbose=False, collect_trajectories=False)
    obs, info = env.reset()
    assert info == {}
    O = _oracle_env(env)
    pos, tgt = _oracle_scene(env, env.scene_generator.last.draws)
    O.reset_envs(np.arange(ne), pos, tgt, np.zeros((4, ne, env.model._execution_spec.tile)))
    np.testing.assert_array_equal(obs.cpu().numpy(), O.observations())
    bufs = C.host_buffers_dict(env.model)
    np.testing.assert_array_equal(bufs["actual_state"][1, :, :n], pos)
    spec = C.spec_from_model(env.model)
    rng = np.random.default_rng(3)
    for k in range(5):
        act = rng.uniform(0, 1, size=(ne, 4)).astype(F32)
        obs, rew, term, tout, info = env.step(torch.from_numpy(act).cuda())
        O.set_action(act, bufs["control_params"])
        C.solve(spec, bufs, kernel_steps=2048, max_steps=10 ** 6)
        r = O.post(bufs["actual_state"], bufs["actual_time"], bufs["time_step"], bufs["actual_event"],
                   bufs["status_flags"], act)
        for name, got in (("terminated", term), ("time_outs", tout)):
            np.testing.assert_array_equal(got.cpu().numpy(), r[name], err_msg=f"{name} step {k}")
        np.testing.assert_allclose(rew.cpu().numpy(), r["rewards"], rtol=0, atol=1e-6, err_msg=f"rewards {k}")
        done = r["done_ids"]
        if done.size:
            np.testing.assert_array_equal(info["dones"].cpu().numpy(), done)
            pos, tgt = _oracle_scene(env, env.scene_generator.last.draws)
            fin = O.reset_envs(done, pos, tgt, bufs["actual_state"])
            np.testing.assert_allclose(info["final_observation"].cpu().numpy(), fin["final_observation"],
                                       rtol=0, atol=1e-6)
            np.testing.assert_allclose(info["episode_return"].cpu().numpy(), fin["episode_return"], rtol=0,
                                       atol=1e-5)
            np.testing.assert_array_equal(info["episode_length"].cpu().numpy(), fin["episode_length"])
            want_obs = fin["observations"]
        else:
            assert info == {}
            want_obs = r["observations"]
        np.testing.assert_allclose(obs.cpu().numpy(), want_obs, rtol=0, atol=1e-6, err_msg=f"observations {k}")
        st = env.model.runtime.device_buffers.state.actual_state.cpu().numpy()[:, :, :n]
        ref = bufs["actual_state"][:, :, :n]
        err = np.abs(st - ref) / (np.abs(ref) + 1e-3)
        assert err.max() < 1e-6, f"state rel err {err.max():.2e} at step {k}"
        print(f"  step {k}: {done.size} done, state err {err.max():.1e}", flush=True)


def test_env_api_behaviour(device, capsys):
    """step(None) samples the action space (same seed, same episode); returned
    tensors stay valid across steps; reset_envs leaves earlier observations alone;
    keep_step keeps the accepted step size; verbose prints the counts."""
    from coupledbubble_control.envs import PosNBC1D

    def make(**kw):
        return PosNBC1D(num_bubbles=2, num_envs=128, envs_per_block=64, time_step_length=1.0, episode_length=2,
                        seed=5, collect_trajectories=False, **kw)

    a, b = make(verbose=False), make(verbose=True)
    oa, _ = a.reset()
    ob, _ = b.reset()
    assert torch.equal(oa, ob)
    ra = a.step(None)
    rb = b.step(None, clone=True)
    for x, y in zip(ra[:4], rb[:4]):
        assert torch.equal(x, y)
    keep = ra[0].clone()
    ra2 = a.step(None)
    assert torch.equal(ra[0], keep), "an earlier step's observations changed"
    ids = ra2[4]["dones"]                                  # episode_length 2: time-outs at least
    assert torch.equal(ids, torch.nonzero(ra2[2] | ra2[3]).flatten())
    assert ra2[3].sum() > 0 and (a.algo_steps[ids] == 0).all()
    assert (a.algo_steps[~(ra2[2] | ra2[3])] > 0).all()
    out = capsys.readouterr().out
    b.step(None)
    out = capsys.readouterr().out
    assert "Number of episode time-outs:" in out
    before = a._observations.clone()
    held = a._observations
    a.reset_envs(torch.arange(10, device="cuda"))
    assert torch.equal(held, before)
    assert not torch.equal(a._observations[:10], before[:10])
    assert torch.equal(a._observations[10:], before[10:])
    c = make(verbose=False, keep_step=True)
    c.reset()
    c.step(None)
    h = c.model.runtime.device_buffers.state.time_step
    assert (h != 1e-6).all(), "keep_step must keep the last accepted step size"
    with pytest.raises(NotImplementedError):
        make(render_env=True)
