"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code (CPU only).

* the C oracle (oracle/ckm1d_oracle.c, in tests/sanitize/oracle_driver.c) on the
  256-system plain workload's first 32 systems (host loop, budget-limited launches),
  the mixed workload's N = 3 and N = 8 bins (uneven horizons, collisions, relaunches)
  and a parameter-study grid with period windows: no sanitizer report, and results
  equal to the regular oracle build's bit for bit;
* the serving ring's host logic (coupled-bubble-control_amd/csrc/cbc_ring.h, in
  tests/sanitize/ring_driver.cpp) against a simulated device over 20 000 rounds, with
  sequence numbers and ring indices wrapping.
Sanitizer reports abort the run (-fno-sanitize-recover=all), so a clean exit means none.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from coupledbubble_control import workloads
from oracle import ckm1d_oracle as O

HERE = Path(__file__).resolve().parent
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc/g++ not available")


def _build(tmp, src, out, cxx):
    cmd = (["g++", "-std=c++17"] if cxx else ["gcc", "-std=c11"]) + SAN + ["-o", str(out), str(src)] + \
        ([] if cxx else ["-lm"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.fixture(scope="module")
def oracle_driver(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("san")
    return _build(tmp, HERE / "sanitize" / "oracle_driver.c", tmp / "oracle_driver", cxx=False)


def _run_case(exe, tmp, w, kernel_steps, periodic=None):
    """Write the case, run the sanitized driver, return its outputs and the regular oracle's."""
    m = workloads.build_model(w)
    b = O.host_buffers_dict(m)
    sp = O.spec_from_model(m)
    ns, ups, tile, nk = sp.ns, sp.ups, sp.tile, sp.nk
    nw = periodic[2] if periodic else 0
    head = np.array([1 if periodic else 0, ns, ups, tile, nk, O.FIELD[sp.field], kernel_steps, nw], np.int32)
    par = np.array([sp.atol, sp.rtol, sp.min_step, sp.max_step, sp.etol, sp.col_threshold,
                    periodic[0] if periodic else 0.0, periodic[1] if periodic else 0.0])
    fin, fout = tmp / "in.bin", tmp / "out.bin"
    with open(fin, "wb") as f:
        head.tofile(f)
        par.tofile(f)
        for k in ("actual_time", "time_end", "time_begin", "time_step", "actual_state", "unit_params",
                  "system_params", "control_params", "coupling_matrices", "actual_event"):
            np.ascontiguousarray(b[k], dtype=np.float64).tofile(f)
    r = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stderr[-3000:])
    raw = np.fromfile(fout, dtype=np.uint8)
    off = 0

    def take(dtype, count):
        nonlocal off
        a = raw[off:off + np.dtype(dtype).itemsize * count].view(dtype)
        off += np.dtype(dtype).itemsize * count
        return a

    got = {"actual_state": take(np.float64, 4 * ns * tile).reshape(4, ns, tile),
           "actual_time": take(np.float64, ns), "time_step": take(np.float64, ns),
           "actual_event": take(np.float64, ns), "status_flags": take(np.int32, ns),
           "total_steps": take(np.int64, ns), "rejected_steps": take(np.int64, ns)}
    if periodic:
        got["max_r"] = take(np.float64, nw * ns * tile).reshape(nw, ns, tile)
        got["mean_x"] = take(np.float64, nw * ns * tile).reshape(nw, ns, tile)
        want = O.solve_periodic(sp, b, periodic[0], int(periodic[1]), nw, max_attempts=1 << 26)
    else:
        want = O.solve(sp, b, kernel_steps=kernel_steps, max_steps=1 << 26)
    want.update({k: b[k] for k in ("actual_state", "actual_time", "time_step", "actual_event", "status_flags")})
    return got, want


def _assert_same(got, want):
    for k, v in got.items():
        np.testing.assert_array_equal(v, want[k], err_msg=k)


def test_oracle_plain_and_relaunches_clean(oracle_driver, tmp_path):
    w = workloads.plain_cpu(seed=0).subset(np.arange(32))
    _assert_same(*_run_case(oracle_driver, tmp_path, w, kernel_steps=500))


@pytest.mark.parametrize("n,k", [(3, 24), (8, 6)])
def test_oracle_mixed_bins_clean(oracle_driver, tmp_path, n, k):
    w = [x for x in workloads.mixed_uneven() if x.n == n][0]
    w = w.subset(np.linspace(0, w.ns - 1, k).astype(int))
    w.t_end = np.minimum(w.t_end, 4.0)
    got, want = _run_case(oracle_driver, tmp_path, w, kernel_steps=700)
    _assert_same(got, want)


def test_oracle_period_windows_clean(oracle_driver, tmp_path):
    w = workloads.param_study(side=4, periods=4.0, discard=1)
    _assert_same(*_run_case(oracle_driver, tmp_path, w, kernel_steps=1 << 20, periodic=(1.0, 1, 3)))


def test_serving_ring_logic_clean(tmp_path):
    exe = _build(tmp_path, HERE / "sanitize" / "ring_driver.cpp", tmp_path / "ring_driver", cxx=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok:"), (r.returncode, r.stderr[-3000:])
