"""cbc_env_post_static (the post-step call without a workspace argument, using the
library's static device workspace) against cbc_env_post: the same outputs bit for bit,
on two calls in a row, the second with fewer envs, so the second call runs on the
workspace the first one left."""
import ctypes

import numpy as np
import pytest
import torch

from test_gpu_env import ENV_PARAMS, _kernels, _model, _parts, _random_inputs, _upload

pytestmark = pytest.mark.gpu


def _pair_step(ne, n, rich, seed):
    rng = np.random.default_rng(seed)
    parts = _parts(n, rich)
    ma, mb = _model(ne, n), _model(ne, n)
    ka, kb = _kernels(ma, parts, **ENV_PARAMS), _kernels(mb, parts, **ENV_PARAMS)
    tile = ma._execution_spec.tile
    xt = rng.uniform(-0.4, 0.4, size=(ne, n))
    for k in (ka, kb):
        k.part_view("XT").copy_(torch.from_numpy(xt.astype(np.float32)))
    state, t, h, ev, status, act = _random_inputs(rng, ne, n, tile, xt)
    for m, k in ((ma, ka), (mb, kb)):
        _upload(m, state, t, h, ev, status)
        k.set_action(torch.from_numpy(act).cuda())
    ca = ka.post(0)
    cb = kb.post(0, static_workspace=True)
    torch.cuda.synchronize()
    assert (ca.done, ca.failures, ca.successes, ca.time_outs) == (cb.done, cb.failures, cb.successes, cb.time_outs)
    assert ca.done > 0
    oa, ob = ka.outputs, kb.outputs
    for name in ("observations", "rewards", "failure", "success", "terminated", "time_outs"):
        assert torch.equal(getattr(oa, name), getattr(ob, name)), name
    assert torch.equal(ka.done_ids[:ca.done], kb.done_ids[:cb.done])
    assert torch.equal(ka.history, kb.history)
    assert torch.equal(ka.total_rewards, kb.total_rewards) and torch.equal(ka.algo_steps, kb.algo_steps)
    da, db = ma.runtime.device_buffers, mb.runtime.device_buffers
    for x, y in ((da.state.actual_time, db.state.actual_time), (da.state.time_step, db.state.time_step),
                 (da.status.actual_event, db.status.actual_event), (da.status.status_flags, db.status.status_flags)):
        assert torch.equal(x, y)
    return ca.done


def test_env_post_static_equals_env_post(device):
    print("case: env_post_static vs env_post, 1000 envs x 3 bubbles then 300 envs x 2", flush=True)
    d1 = _pair_step(1000, 3, True, seed=71)
    d2 = _pair_step(300, 2, False, seed=72)
    print(f"  done lists {d1}, {d2}", flush=True)


def test_env_post_static_rejects_too_many_envs(device):
    from coupledbubble_control.backends.hip import _abi

    print("case: env_post_static size limit", flush=True)
    parts = _parts(2, False)
    k = _kernels(_model(64, 2), parts, **ENV_PARAMS)
    spec = _abi.EnvSpecC.from_buffer_copy(k.spec)
    spec.num_envs = _abi.CBC_ENV_STATIC_MAX_ENVS + 1
    b = k._buffers(None)
    rc = _abi.lib().cbc_env_post_static(ctypes.byref(spec), ctypes.byref(b), 0, _abi.stream_handle(None), None)
    assert rc == _abi.CBC_ERR_INVALID
