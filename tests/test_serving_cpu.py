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
