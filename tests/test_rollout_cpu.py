"""Host logic of the overlapped rollout (no GPU): groups that finish in random order,
and publishing held back while the ring is full, give every env the observations,
actions and failures of the synchronous loop."""
import numpy as np

from coupledbubble_control.rollout import OverlappedRollout


class _Queue:
    def __init__(self, env):
        self.env = env

    def poll(self):
        self.env.tick()
        return 0


class FakeServedEnv:
    """Deterministic per-env dynamics; a published step finishes after a random number of
    polls, and submit refuses (ring full) at random."""

    def __init__(self, ne=37, nb=2, nk=2, seed=0, refuse=0.3):
        self.num_envs, self.num_bubbles, self.nk = ne, nb, nk
        self.state = np.linspace(0.1, 0.9, ne * nb).reshape(ne, nb)
        self.rng = np.random.default_rng(seed)
        self.refuse = refuse
        self.pending = {}          # seq -> [ids, actions, polls left]
        self.done_seqs = set()
        self.results = {}
        self.ids_of = {}
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
        s = self.next_seq
        self.next_seq += 1
        self.ids_of[s] = np.asarray(ids).copy()
        self.pending[s] = [np.asarray(ids).copy(), np.asarray(a).copy(), int(self.rng.integers(0, 4))]
        return s

    def tick(self):
        for s, p in list(self.pending.items()):
            if p[2] == 0:
                self.results[s] = self._advance(p[0], p[1])
                self.done_seqs.add(s)
                del self.pending[s]
            else:
                p[2] -= 1

    def finished(self, ids, seq):
        return seq in self.done_seqs

    def collect(self, ids, rng=None):
        s = max(k for k in self.done_seqs if np.array_equal(self.ids_of[k], ids))
        return self.results[s]

    def step_sync(self, a):
        return self._advance(np.arange(self.num_envs), a)


def _policy(o, ids):
    return np.stack([0.5 + 0.4 * np.cos(7.0 * o[:, 0] + ids), 0.3 + 0.2 * o[:, 1]], axis=1)


def test_grouped_rollout_equals_synchronous_loop():
    steps = 6
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
    for groups, seed in ((1, 1), (3, 2), (5, 3), (37, 4)):
        env = FakeServedEnv(seed=seed)
        r = OverlappedRollout(env, _policy, groups=groups).collect(obs0, steps)
        np.testing.assert_array_equal(r["observations"], np.stack(obs))
        np.testing.assert_array_equal(r["actions"], np.stack(acts))
        np.testing.assert_array_equal(r["failures"], np.stack(fails))
        assert (r["group_done_s"] > 0).all()


def test_groups_out_of_range():
    import pytest

    env = FakeServedEnv(ne=4)
    with pytest.raises(ValueError):
        OverlappedRollout(env, _policy, groups=0)
    with pytest.raises(ValueError):
        OverlappedRollout(env, _policy, groups=5)
