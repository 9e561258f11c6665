"""ServedPosNBC1D: the PosNBC1D step served block by block (envs/served.py).

* one block: a served rollout equals PosNBC1D.step in a loop bit for bit
  (observations, rewards, terminations, time-outs, done lists);
* three blocks stepping on their own equal the same three blocks in lockstep bit for
  bit (the result does not depend on which block finishes first);
* three blocks against the oracle environment (oracle/env_step.py, one per block)
  driven by the C oracle solver with the same actions and scene draws: observations
  and rewards within 1e-6, flags and done lists exactly, states within 1e-6.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32 = np.float32


def _env(n=2, ne=96, seed=17):
    from coupledbubble_control.envs import PosNBC1D

    return PosNBC1D(num_bubbles=n, num_envs=ne, envs_per_block=64, R0=[40.0, 60.0, 50.0][:n], PA=[1.0, 0.5],
                    time_step_length=2.0, episode_length=3, position_tolerance=0.05, seed=seed, verbose=False,
                    collect_trajectories=False)


def _policy(obs, begin):
    """A row-wise deterministic policy (elementwise float32 arithmetic only, so a row's
    actions do not depend on the batch it is in): four actions in [0.05, 0.95]."""
    d = obs.shape[1]
    cols = [0.5 + 0.45 * torch.sin(obs[:, j % d] * (1.3 + j) + obs[:, (j + 1) % d]) for j in range(4)]
    return torch.stack(cols, dim=1).to(torch.float32)


def test_one_block_equals_posnbc1d_steps(device):
    from coupledbubble_control.envs import ServedPosNBC1D

    print("case: served PosNBC1D, 1 block, 96 envs x 2 bubbles, 6 steps vs PosNBC1D.step", flush=True)
    a, b = _env(), _env()
    oa, _ = a.reset()
    ob, _ = b.reset()
    assert torch.equal(oa, ob)
    served = ServedPosNBC1D(b, groups=1)
    try:
        r = served.rollout(_policy, steps=6)
    finally:
        served.close()
    obs = oa
    ndone = 0
    for k in range(6):
        act = _policy(obs, 0)
        obs, rew, term, tout, info = a.step(act, clone=True)
        assert torch.equal(r["actions"][k], act.cpu()), f"actions {k}"
        assert torch.equal(r["observations"][k + 1], obs.cpu()), f"observations {k}"
        assert torch.equal(r["rewards"][k], rew.cpu()), f"rewards {k}"
        assert torch.equal(r["terminated"][k], term.cpu()) and torch.equal(r["time_outs"][k], tout.cpu())
        if info:
            ids, fin = r["dones"][(k, 0)]
            assert torch.equal(ids, info["dones"].cpu())
            assert torch.equal(fin[0], info["final_observation"].cpu())
            assert torch.equal(fin[1], info["episode_return"].cpu())
            assert torch.equal(fin[2], info["episode_length"].cpu())
            ndone += ids.numel()
        else:
            assert (k, 0) not in r["dones"]
    assert ndone > 0
    print(f"  {ndone} episodes ended over 6 steps", flush=True)


def test_blocks_on_their_own_equal_lockstep(device):
    from coupledbubble_control.envs import ServedPosNBC1D

    print("case: served PosNBC1D, 3 blocks overlapped vs lockstep, 96 envs x 3 bubbles, 5 steps", flush=True)
    res = []
    for lockstep in (False, True):
        e = _env(n=3)
        e.reset()
        s = ServedPosNBC1D(e, groups=3)
        try:
            res.append(s.rollout(_policy, steps=5, lockstep=lockstep))
        finally:
            s.close()
    for key in ("observations", "actions", "rewards", "terminated", "time_outs"):
        assert torch.equal(res[0][key], res[1][key]), key
    assert res[0]["dones"].keys() == res[1]["dones"].keys()
    for k in res[0]["dones"]:
        assert torch.equal(res[0]["dones"][k][0], res[1]["dones"][k][0])


def _slice(bufs, lo, hi):
    """The reference-layout buffers of systems lo..hi-1 (the system axis is the one of
    length NS: axis 0 of (NS,) rows, else the axis after the leading row axes)."""
    out = {}
    for k, v in bufs.items():
        if v.ndim == 1:
            out[k] = np.ascontiguousarray(v[lo:hi])
        elif k == "coupling_matrices":
            out[k] = np.ascontiguousarray(v[:, :, lo:hi])
        elif k == "dense_state":
            out[k] = np.ascontiguousarray(v[:, :, lo:hi])
        else:
            out[k] = np.ascontiguousarray(v[:, lo:hi])
    return out


def test_blocks_match_oracle_environment(device):
    from test_gpu_env import _oracle_env

    from coupledbubble_control.envs import ServedPosNBC1D
    from oracle import ckm1d_oracle as C
    from oracle import env_step as E

    print("case: served PosNBC1D, 3 blocks vs oracle environment, 96 envs x 2 bubbles, 5 steps", flush=True)
    n, ne, steps = 2, 96, 5
    env = _env(n=n, ne=ne, seed=23)
    obs0, _ = env.reset()
    bufs_all = C.host_buffers_dict(env.model)
    d0 = {k: v.cpu().numpy() for k, v in env.scene_generator.last.draws.items()}
    served = ServedPosNBC1D(env, groups=3)
    served.draw_log = []
    try:
        r = served.rollout(_policy, steps=steps)
    finally:
        served.close()
    g = env.scene_generator
    sp = E.SceneParams(num_bubbles=n, domain_min=g.domain_min, domain_max=g.domain_max, margin=g.margin,
                       min_distance=g.min_distance, max_distance=g.max_distance, alignment=g.alignment)

    def scene(d):
        return (E.random_uniform_positions(sp, d["bubbles_gaps"], d.get("bubbles_align")),
                E.random_uniform_positions(sp, d["targets_gaps"], d.get("targets_align")))

    log = {(gg, st): (ids, d) for gg, st, ids, d in served.draw_log}
    pos0, tgt0 = scene(d0)
    spec_all = C.spec_from_model(env.model)
    total_done = 0
    for gi, (b, c) in enumerate(served.blocks):
        O = _oracle_env(env, num_envs=c)
        bufs = _slice(bufs_all, b, b + c)
        spec = C.OracleSpec(**{**spec_all.__dict__, "ns": c})
        O.reset_envs(np.arange(c), pos0[b:b + c], tgt0[b:b + c], np.zeros_like(bufs["actual_state"]))
        np.testing.assert_array_equal(r["observations"][0, b:b + c].numpy(), O.observations())
        for k in range(steps):
            act = r["actions"][k, b:b + c].numpy().astype(F32)
            O.set_action(act, bufs["control_params"])
            C.solve(spec, bufs, kernel_steps=2048, max_steps=10 ** 6)
            o = O.post(bufs["actual_state"], bufs["actual_time"], bufs["time_step"], bufs["actual_event"],
                       bufs["status_flags"], act)
            for name in ("terminated", "time_outs"):
                np.testing.assert_array_equal(r[name][k, b:b + c].numpy(), o[name], err_msg=f"{name} {gi}/{k}")
            np.testing.assert_allclose(r["rewards"][k, b:b + c].numpy(), o["rewards"], rtol=0, atol=1e-6)
            done = o["done_ids"]
            want_obs = o["observations"]
            if done.size:
                ids, d = log[(gi, k)]
                np.testing.assert_array_equal(ids, done + b)
                pos, tgt = scene(d)
                fin = O.reset_envs(done, pos, tgt, bufs["actual_state"])
                want_obs = fin["observations"]
                total_done += done.size
            else:
                assert (gi, k) not in log
            np.testing.assert_allclose(r["observations"][k + 1, b:b + c].numpy(), want_obs, rtol=0, atol=1e-6)
        st = env.model.runtime.device_buffers.state.actual_state.cpu().numpy()[:, b:b + c, :n]
        ref = bufs["actual_state"][:, :, :n]
        err = np.abs(st - ref) / (np.abs(ref) + 1e-3)
        assert err.max() < 1e-6, f"block {gi}: state rel err {err.max():.2e}"
    assert total_done > 0
    print(f"  {total_done} episodes ended", flush=True)
