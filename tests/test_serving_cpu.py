"""Host logic of the serving loop (no GPU): admission under the in-flight budget,
per-vector-env ordering, the vector-env partition, the Poisson request stream,
the per-system parameter rebuild used by episode resets, and argument checks of
the serving C ABI that fail before any device call."""
import ctypes

import numpy as np
import pytest

from coupledbubble_control.serving import Admission, Request, inflight_bytes, make_requests, vector_envs


def _req(rid, venv, size, n=2):
    return Request(rid, n, venv, np.arange(size, dtype=np.int32), np.zeros((2, size)), 0.0)


def test_admission_budget_and_order():
    a = Admission(max_inflight=100)
    for r in (_req(0, 0, 60), _req(1, 1, 30), _req(2, 2, 50), _req(3, 3, 5)):
        a.push(r)
    got = [r.rid for r in a.admit()]
    # 60 + 30 fit; 50 does not and holds back the 5 behind it (no overtaking on budget)
    assert got == [0, 1]
    assert a.inflight == 90
    assert [r.rid for r in a.pending] == [2, 3]
    a.finish(_req(0, 0, 60))
    assert [r.rid for r in a.admit()] == [2, 3]
    assert a.inflight == 85


def test_admission_one_step_per_vector_env():
    a = Admission(max_inflight=1000)
    a.push(_req(0, 7, 10))
    a.push(_req(1, 7, 10))     # same vector env: waits for step 0
    a.push(_req(2, 8, 10))     # other env may pass it
    assert [r.rid for r in a.admit()] == [0, 2]
    assert [r.rid for r in a.admit()] == []
    a.finish(_req(0, 7, 10))
    assert [r.rid for r in a.admit()] == [1]


def test_admission_by_bin_and_oversize():
    a = Admission(max_inflight=64)
    with pytest.raises(ValueError):
        a.push(_req(0, 0, 65))
    a.push(_req(1, 1, 10, n=4))
    a.push(_req(2, 2, 10, n=2))
    assert [r.rid for r in a.admit(only_bin=2)] == [2]
    assert [r.rid for r in a.admit(only_bin=4)] == [1]


def test_inflight_bytes():
    assert inflight_bytes(2, 2) == 48 + 32 + 64
    assert inflight_bytes(3, 4) == 48 + 32 + 128


def test_vector_env_partition():
    rng = np.random.default_rng(0)
    v = vector_envs({2: 12288, 4: 4096}, rng)
    for n, ns in ((2, 12288), (4, 4096)):
        ids = np.concatenate([i for m, i in v if m == n])
        np.testing.assert_array_equal(np.sort(ids), np.arange(ns))
        sizes = [len(i) for m, i in v if m == n]
        assert min(sizes) >= 64 and max(sizes) <= 1024 + 64


def test_vector_env_partition_small_bins():
    for seed in range(20):
        v = vector_envs({2: 384, 4: 192}, np.random.default_rng(seed), lo=64, hi=160)
        for n, ns in ((2, 384), (4, 192)):
            ids = np.concatenate([i for m, i in v if m == n])
            np.testing.assert_array_equal(np.sort(ids), np.arange(ns))
            assert min(len(i) for m, i in v if m == n) >= 64


def test_poisson_requests():
    rng = np.random.default_rng(1)
    v = vector_envs({2: 12288, 4: 4096}, rng)
    reqs = make_requests(v, rate_env_steps=200000.0, duration=2.0, seed=3)
    assert len(make_requests(v, rate_env_steps=200000.0, count=17, seed=3)) == 17
    env_steps = sum(r.size for r in reqs)
    assert 0.8 * 400000 < env_steps < 1.2 * 400000
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0) and t[-1] <= 2.0
    # round robin: each vector env appears once before any appears twice
    first = [r.venv for r in reqs[:len(v)]]
    assert len(set(first)) == len(v)
    for r in reqs[:5]:
        assert r.actions.shape == (2, r.size) and (r.actions >= 0).all() and (r.actions <= 1.5).all()


def test_reset_parameter_rows_match_full_build():
    """The per-system rebuild used by reset_systems gives the rows of the full build."""
    from coupledbubble_control import workloads

    for n in (2, 4):
        w = workloads.serving_bins(seed=2, sizes={n: 32})[n]
        m = workloads.build_model(w)
        up, sp, cm = m.build_parameters()
        idx = np.array([3, 17, 30])
        up_s, cm_s = m._unit_and_coupling(m._R0[idx])
        np.testing.assert_array_equal(up_s, up[:, idx])
        np.testing.assert_array_equal(cm_s, cm[:, :, idx])


def test_serve_abi_rejects_null_handles():
    from coupledbubble_control.backends.hip import _abi

    lib = _abi.lib()
    spec = _abi.ServeSpecC(64, 0, 1, 8, 1000, 1000, 1)
    h = ctypes.c_void_p()
    assert lib.cbc_serve_create(None, None, ctypes.byref(spec), None, ctypes.byref(h)) == _abi.CBC_ERR_INVALID
    assert lib.cbc_serve_submit(None, None, None, 0, 1.0, None, None, None) == _abi.CBC_ERR_INVALID
    assert lib.cbc_serve_poll(None, None) == _abi.CBC_ERR_INVALID
    assert lib.cbc_serve_destroy(None) == _abi.CBC_OK


# -- the request loop's per-env ordering, driven by a host-side stand-in for the serving queue


class _FakeQueue:
    """Host stand-in for ServeQueue: an item of env e finishes ``dur[e]`` polls after it is
    published; every submission is logged as (tick, env, first amplitude)."""

    def __init__(self, ns, dur, refuse_at=()):
        self.dur = dict(dur)
        self.tick = 0
        self.seq = 0
        self.fin_tick = {}
        self.log = []
        self.refuse_at = set(refuse_at)
        self.calls = 0
        self.done_seq = np.zeros(ns, dtype=np.int64) - 1
        self.out_misc = np.zeros((ns, 4))
        self.live = {}

    def submit(self, ids, pa, t_end, h0=0.0, reset=None):
        self.calls += 1
        if self.calls in self.refuse_at:
            return None
        seq = np.arange(self.seq, self.seq + ids.size, dtype=np.uint32)
        self.seq += ids.size
        for e, s, a in zip(ids.tolist(), seq.tolist(), pa[0].tolist()):
            assert e not in self.live, f"env {e} published while its previous item runs"
            self.live[e] = s
            self.fin_tick[s] = self.tick + self.dur.get(e, 1)
            self.log.append((self.tick, e, a))
        return seq

    def poll(self):
        self.tick += 1
        for e, s in list(self.live.items()):
            if self.fin_tick[s] <= self.tick:
                self.out_misc[e] = (5.0, 1e-3, 1.0, 10.0)
                self.done_seq[e] = s
                del self.live[e]
        return len(self.live)

    def done(self, ids, seq):
        return self.done_seq[np.asarray(ids)] >= np.asarray(seq, dtype=np.int64)

    def results(self, ids):
        return None, self.out_misc[np.asarray(ids)].copy()

    def stats(self):
        return (1, self.seq, self.seq)

    def close(self):
        pass


def _host_engine(queue, ns, order, budget=10 ** 6):
    from types import SimpleNamespace

    from coupledbubble_control.serving import ServingEngine

    class HostEngine(ServingEngine):
        def __init__(self):
            spec = SimpleNamespace(ns=ns)
            self.bins = {2: SimpleNamespace(_execution_spec=spec, _model_spec=SimpleNamespace(nk=2))}
            self.mode, self.order, self.periods, self.keep_step = "continuous", order, 5.0, True
            self.reset_prob, self.reset_scene, self.rng = 0.0, None, np.random.default_rng(0)
            self.admission = Admission(budget, per_venv=order == "venv")
            self.queues = {2: queue}
            self.pending_reset = {2: np.zeros(ns, dtype=bool)}
            self._init_bookkeeping({2: ns})

    return HostEngine()


def _stream(steps=3):
    """Two vector envs of three envs each, ``steps`` requests each, all arriving at t = 0;
    request k of venv v carries amplitude 10 v + k for every env."""
    reqs = []
    for k in range(steps):
        for v in range(2):
            ids = np.arange(3 * v, 3 * v + 3, dtype=np.int32)
            reqs.append(Request(len(reqs), 2, v, ids, np.full((2, 3), 10.0 * v + k), 0.0))
    return reqs


def _per_env_actions(log):
    out = {}
    for _, e, a in log:
        out.setdefault(e, []).append(a / 1e5)
    return out


@pytest.mark.parametrize("order", ["env", "venv"])
def test_request_loop_orders_steps_per_env(order):
    # env 1 is slow: 7 polls per step, the others 1
    q = _FakeQueue(6, {1: 7})
    eng = _host_engine(q, 6, order)
    st = eng.run(_stream(), timeout_s=30.0)
    assert st["requests"] == 6 and st["env_steps"] == 18
    # every env ran its three steps in request order
    for e, acts in _per_env_actions(q.log).items():
        v = e // 3
        assert acts == [10.0 * v + k for k in range(3)], (e, acts)
    for r in eng.completed:
        assert r.flags.tolist() == [1, 1, 1] and r.periods == 15.0
    slow_done = min(t for t, e, a in q.log if e == 1 and a == 1e5)   # env 1's 2nd step published
    others_2nd = [t for t, e, a in q.log if e in (0, 2) and a == 1e5]
    if order == "env":
        # envs 0 and 2 start step 2 long before the slow env 1 has finished step 1
        assert max(others_2nd) < slow_done
        # requests still complete in order within a vector env
        rids = [r.rid for r in eng.completed if r.venv == 0]
        assert rids == sorted(rids)
    else:
        assert min(others_2nd) >= slow_done
    assert eng.admission.inflight == 0 and not eng.live


def test_request_loop_env_order_budget_and_full_ring():
    # budget of 4 in-flight env steps: requests of 3 envs are admitted one at a time
    # (3 + 3 > 4); the ring refuses the 2nd and 3rd submissions, which are published later
    q = _FakeQueue(6, {1: 3, 4: 5}, refuse_at=(2, 3))
    eng = _host_engine(q, 6, "env", budget=4)
    peak = []
    orig = eng._publish

    def publish(now):
        peak.append(eng.admission.inflight)
        orig(now)

    eng._publish = publish
    st = eng.run(_stream(), timeout_s=30.0)
    assert st["requests"] == 6 and st["ring_full_retries"] == 2
    assert max(peak) <= 4
    for e, acts in _per_env_actions(q.log).items():
        assert acts == [10.0 * (e // 3) + k for k in range(3)]
