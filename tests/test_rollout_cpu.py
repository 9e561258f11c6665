"""Host logic of the overlapped rollout (no GPU): groups, or single envs, that finish in
random order, and publishing held back while the ring is full, give every env the
observations, actions and failures of the synchronous loop."""
import numpy as np

from coupledbubble_control.rollout import OverlappedRollout


class _Queue:
    def __init__(self, env):
        self.env = env

    def poll(self):
        self.env.tick()
        return 0


class FakeServedEnv:
    """Deterministic per-env dynamics; a published item finishes after a random number of
    polls, and submit refuses (ring full) at random.  As on the serving queue, every item
    has its own sequence number and each env keeps the result of its last finished item."""

    def __init__(self, ne=37, nb=2, nk=2, seed=0, refuse=0.3):
        self.num_envs, self.num_bubbles, self.nk = ne, nb, nk
        self.state = np.linspace(0.1, 0.9, ne * nb).reshape(ne, nb)
        self.rng = np.random.default_rng(seed)
        self.refuse = refuse
        self.pending = []          # [env id, action row, seq, polls left]
        self.done_seq = np.zeros(ne, dtype=np.int64)
        self.obs = np.zeros((ne, nb), dtype=np.float32)
        self.fail = np.zeros(ne, dtype=bool)
        self.next_seq = 1
        self.queue = _Queue(self)

    def _advance(self, ids, a):
        x = self.state[ids]
        x = np.mod(x + 0.1 * np.sin(3.0 * x) + 0.05 * a[:, :1], 1.0)
        self.state[ids] = x
        return x.astype(np.float32), x[:, 0] > 0.8

    def submit(self, ids, a):
        if self.rng.random() < self.refuse:
            return None
        ids = np.asarray(ids)
        seq = np.arange(self.next_seq, self.next_seq + ids.size)
        self.next_seq += ids.size
        for e, row, s in zip(ids, np.asarray(a), seq):
            self.pending.append([int(e), row.copy(), int(s), int(self.rng.integers(0, 4))])
        return seq

    def tick(self):
        keep = []
        for p in self.pending:
            if p[3] == 0:
                o, f = self._advance(np.array([p[0]]), p[1][None, :])
                self.obs[p[0]], self.fail[p[0]], self.done_seq[p[0]] = o[0], f[0], p[2]
            else:
                p[3] -= 1
                keep.append(p)
        self.pending = keep

    def done_mask(self, ids, seq):
        return self.done_seq[np.asarray(ids)] >= np.asarray(seq)

    def finished(self, ids, seq):
        return bool(self.done_mask(ids, seq).all())

    def collect(self, ids, rng=None, record=True):
        ids = np.asarray(ids)
        return self.obs[ids].copy(), self.fail[ids].copy()

    def step_sync(self, a):
        return self._advance(np.arange(self.num_envs), a)


def _policy(o, ids):
    return np.stack([0.5 + 0.4 * np.cos(7.0 * o[:, 0] + ids), 0.3 + 0.2 * o[:, 1]], axis=1)


def _synchronous(steps):
    ref_env = FakeServedEnv()
    obs0 = ref_env.state.astype(np.float32).copy()
    obs, acts, fails = [obs0], [], []
    ids = np.arange(ref_env.num_envs)
    for _ in range(steps):
        a = _policy(obs[-1], ids)
        o, f = ref_env.step_sync(a)
        obs.append(o)
        acts.append(a)
        fails.append(f)
    return obs0, np.stack(obs), np.stack(acts), np.stack(fails)


def _check(r, obs, acts, fails):
    np.testing.assert_array_equal(r["observations"], obs)
    np.testing.assert_array_equal(r["actions"], acts)
    np.testing.assert_array_equal(r["failures"], fails)
    assert (r["group_done_s"] > 0).all()


def test_grouped_rollout_equals_synchronous_loop():
    steps = 6
    obs0, obs, acts, fails = _synchronous(steps)
    for groups, seed in ((1, 1), (3, 2), (5, 3), (37, 4)):
        env = FakeServedEnv(seed=seed)
        _check(OverlappedRollout(env, _policy, groups=groups).collect(obs0, steps), obs, acts, fails)


def test_per_env_rollout_equals_synchronous_loop():
    """Every env on its own: batches of the envs that finished (at least min_batch of
    them, or those that waited max_wait_s), in random completion order, with refusals."""
    steps = 6
    obs0, obs, acts, fails = _synchronous(steps)
    for min_batch, wait, seed in ((1, 0.0, 5), (5, 1e-3, 6), (37, 10.0, 7), (None, 1e-4, 8)):
        env = FakeServedEnv(seed=seed)
        r = OverlappedRollout(env, _policy, groups="env", min_batch=min_batch, max_wait_s=wait).collect(obs0, steps)
        _check(r, obs, acts, fails)
        assert r["env_done_s"].shape == (env.num_envs,) and (r["env_done_s"] > 0).all()


def test_groups_out_of_range():
    import pytest

    env = FakeServedEnv(ne=4)
    for bad in (0, 5, "envs", 1.5):
        with pytest.raises(ValueError):
            OverlappedRollout(env, _policy, groups=bad)
    with pytest.raises(ValueError):
        OverlappedRollout(env, _policy, groups="env", min_batch=0)
