"""Trajectory collection on the GPU (cbc_traj_append / cbc_traj_gather behind
TrajectorBuffer) against the oracle recorder (oracle/trajectory.py), read back
from the zarr files the writer produced.

Kernel level: random dense buffers with ragged point counts (0, 1, ..., NDO per
env and step), several steps, episodes ending for a subset of envs at different
steps -- every written array equals the oracle's merged float32 array bit for
bit (both round the same float64 values once; times are t + k * 5.0 in float64
before rounding).  Environment level: PosNBC1D with collect_trajectories
against the oracle environment and recorder driven by the C oracle solver with
dense output; observations, rewards and states within the parity bar, sample
times to the dense-output bar, point counts and episode metadata exactly.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32 = np.float32


def _read_episode(root, eid):
    from coupledbubble_control.envs.common import zarr_v3 as Z

    start = (eid // 1000) * 1000
    g = root / f"shard_{start:05d}_{start + 999:05d}.zarr" / f"trj_{eid:06d}"
    out = {name: Z.read_array(g / name) for name in Z.children(g)}
    out["attrs"] = Z.group_attrs(g)
    return out


def test_append_and_gather_match_oracle_recorder(device, tmp_path):
    from coupledbubble_control.envs.common.trajectory_buffer import TrajectorBuffer
    from oracle.trajectory import TrajectoryRecorder

    print("case: trajectory kernels, 37 envs x 3 bubbles, ndo 6", flush=True)
    ne, n, tile, ndo, obs_dim, na = 37, 3, 4, 6, 9, 4
    rng = np.random.default_rng(2)
    buf = TrajectorBuffer(ne, n, saved_states=[0, 1, 3], zarr_root=str(tmp_path / "store"), max_episode_steps=5,
                          verbose=False)
    rec = TrajectoryRecorder(ne, n, saved_states=(0, 1, 3))
    ended = 0
    for step in range(7):
        obs = rng.random((ne, obs_dim)).astype(F32)
        act = rng.random((ne, na)).astype(F32)
        rew = rng.normal(size=ne).astype(F32)
        di = rng.integers(0, ndo + 1, size=ne).astype(np.int32)
        dt = rng.random((ndo, ne)) * 5.0
        ds = rng.normal(size=(ndo, 4, ne, tile))
        d_di = torch.from_numpy(di).cuda()
        buf.step(torch.from_numpy(obs).cuda(), torch.from_numpy(act).cuda(), torch.from_numpy(rew).cuda(), d_di,
                 torch.from_numpy(dt).cuda(), torch.from_numpy(ds).cuda())
        rec.step(obs, act, rew, di, dt, ds)
        assert (d_di == 0).all(), "dense_index must be cleared"
        ids = np.sort(rng.choice(ne, size=rng.integers(0, 12), replace=False)) if step in (2, 4, 6) else []
        if step == 4:
            ids = np.arange(ne)                     # every env ends at once
        if len(ids):
            fo = rng.random((len(ids), obs_dim)).astype(F32)
            ln = rng.integers(1, 9, size=len(ids))
            rt = rng.normal(size=len(ids)).astype(F32)
            radii = rng.uniform(20, 80, size=(len(ids), n))
            buf.end_episode(torch.from_numpy(ids), torch.from_numpy(fo), torch.from_numpy(ln), torch.from_numpy(rt),
                            radii)
            rec.end_episode(ids, fo, ln, rt, radii)
            ended += len(ids)
    assert buf.episode_count == ended == len(rec.episodes)
    for eid, want in enumerate(rec.episodes):
        got = _read_episode(tmp_path / "store", eid)
        for k in ("observations", "actions", "rewards", "dense_time", "equilibrium_radii", "dense_state_0",
                  "dense_state_1", "dense_state_3"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"episode {eid} {k}")
        a = got["attrs"]
        assert a["env_id"] == eid and a["num_units"] == n
        assert a["episode_length"] == want["episode_length"] and a["episode_reward"] == want["episode_reward"]


def test_env_trajectories_match_oracle(device, tmp_path):
    from coupledbubble_control.envs import PosNBC1D
    from coupledbubble_control.envs.common import zarr_v3 as Z
    from oracle import ckm1d_oracle as C
    from oracle.trajectory import TrajectoryRecorder
    from test_gpu_env import _oracle_env, _oracle_scene
    from test_gpu_plain import ATOL_FLOOR, RTOL

    print("case: PosNBC1D trajectories, 16 envs x 2 bubbles, ndo 8", flush=True)
    ne, n = 16, 2
    root = tmp_path / "PosNBC1D"
    env = PosNBC1D(num_bubbles=n, num_envs=ne, envs_per_block=64, R0=[40.0, 60.0], PA=[1.0, 0.5],
                   time_step_length=1.0, episode_length=3, position_tolerance=0.05, seed=23, verbose=False,
                   collect_trajectories=True, trajectory_resolution=8, trajectory_buffer_store=str(root))
    env.reset()
    O = _oracle_env(env)
    pos, tgt = _oracle_scene(env, env.scene_generator.last.draws)
    O.reset_envs(np.arange(ne), pos, tgt, np.zeros((4, ne, 2)))
    rec = TrajectoryRecorder(ne, n)
    bufs = C.host_buffers_dict(env.model)
    spec = C.spec_from_model(env.model)
    rng = np.random.default_rng(4)
    for k in range(5):
        act = rng.uniform(0, 1, size=(ne, 4)).astype(F32)
        env.step(torch.from_numpy(act).cuda())
        O.set_action(act, bufs["control_params"])
        C.solve(spec, bufs, kernel_steps=2048, max_steps=10 ** 6)
        r = O.post(bufs["actual_state"], bufs["actual_time"], bufs["time_step"], bufs["actual_event"],
                   bufs["status_flags"], act)
        rec.step(r["observations"], act, r["rewards"], bufs["dense_index"], bufs["dense_time"], bufs["dense_state"])
        done = r["done_ids"]
        if done.size:
            pos, tgt = _oracle_scene(env, env.scene_generator.last.draws)
            fin = O.reset_envs(done, pos, tgt, bufs["actual_state"])
            rec.end_episode(done, fin["final_observation"], fin["episode_length"], fin["episode_return"],
                            np.tile([40.0, 60.0], (done.size, 1)))
    assert env.trajectory_buffer.episode_count == len(rec.episodes) > 0
    meta = Z.group_attrs(root / "global_metadata.zarr")
    assert meta["num_bubbles"] == n and meta["time_step_length"] == 1.0 and meta["_initialized"] is True
    for eid, want in enumerate(rec.episodes):
        got = _read_episode(root, eid)
        assert got["attrs"]["episode_length"] == want["episode_length"]
        np.testing.assert_array_equal(got["actions"], want["actions"])
        np.testing.assert_array_equal(got["equilibrium_radii"], want["equilibrium_radii"])
        np.testing.assert_allclose(got["observations"], want["observations"], rtol=0, atol=1e-6)
        np.testing.assert_allclose(got["rewards"], want["rewards"], rtol=0, atol=1e-6)
        assert got["dense_time"].shape == want["dense_time"].shape, f"episode {eid} point count"
        np.testing.assert_allclose(got["dense_time"], want["dense_time"], rtol=1e-6)
        for r_ in (0, 1):
            g, w = got[f"dense_state_{r_}"], want[f"dense_state_{r_}"]
            scale = np.abs(w).max() + ATOL_FLOOR
            np.testing.assert_allclose(g, w, rtol=0, atol=max(RTOL * scale, 2e-6 * scale))
