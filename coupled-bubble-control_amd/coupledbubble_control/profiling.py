"""Measurement helpers for profiling runs (``python -m coupledbubble_control.profiling``).

* ``copy``     -- a plain device copy of known size (default 512 MiB read and
                  written per launch, 3 launches): under ``rocprofv3 --pmc
                  FETCH_SIZE`` / ``WRITE_SIZE`` it calibrates the counters'
                  scale against known bytes.
* ``latency``  -- critical-path time per Cash--Karp attempt of every mapping
                  and bubble count: one plain launch integrating one period of
                  a serving-style ensemble, kernel time (HIP events, best of 3)
                  over the largest attempt count of any system.  Prints one
                  JSON object per case; ``--systems`` reruns every case at the
                  given ensemble sizes (occupancy sweep).
* ``param``    -- ``--steps`` launches of the parameter-study kernel
                  (65 536 systems of N=2, 10 periods each), the dominant
                  kernel of that workload, for counter collection.
* ``sweep``    -- the reference's kernel benchmark sweep
                  (scripts/benchmarks/prof_simu_time.sh: NS = 256 .. 32 768,
                  seeds 11 42 55 77 89 119, kernel_steps 512, one warm-up and 5
                  measured solves of 5 periods from the initial state) on its
                  scenes (``--scenes``, default 2B_TYPICAL): per NS the mean
                  solve time, sim-time per wall-second (kernel_prof.py:246-252),
                  ensemble periods/s and the FP64 roofline fraction; one JSON
                  object per NS.
"""
from __future__ import annotations

import argparse
import json
import sys
from typing import Dict, Iterable, List, Sequence, Tuple

import numpy as np

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) peak


def flops_per_attempt(n: int) -> int:
    """Algorithmic FLOPs of one Cash--Karp attempt of an N-bubble system (NK=2):
    6 RHS evaluations, each 129 N (uncoupled, field) + 49 N(N-1) (explicit
    coupling and matrix assembly per ordered pair) + the 2N x 2N LU solve
    (2/3 m^3 + 3/2 m^2 - 7/6 m, m = 2N); plus 228 N for stage sums, the 5th-order
    solution and the error norm, and 17 for stage times and the controller.
    Each + - x / and each sin/cos/pow counts as one."""
    m = 2 * n
    solve = (2 * m ** 3) / 3 + 1.5 * m ** 2 - 7 * m / 6
    rhs = 129 * n + 49 * n * (n - 1) + solve
    return int(round(6 * rhs + 228 * n + 17))


LATENCY_CASES: Tuple[Tuple[int, str, int], ...] = (
    (2, "lane", 4096), (2, "group", 4096), (3, "group", 4096), (4, "group", 4096), (4, "lane", 4096),
    (8, "group", 2048), (16, "group", 1024))


def copy_kernel(mib: int = 512, reps: int = 3) -> int:
    """Copy a float64 buffer of `mib` MiB `reps` times; returns the bytes per launch."""
    import torch

    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return n * 8


def attempt_latency(cases: Iterable[Tuple[int, str, int]] = LATENCY_CASES, periods: float = 1.0,
                    reps: int = 3, event_mode: str = "bracket") -> List[Dict]:
    """Critical-path microseconds per attempt for (bubbles, mapping, systems) cases, in the
    given collision-event mode ("exact" with etol 1e-12)."""
    import dataclasses

    import torch

    from . import workloads as W
    from .backends.hip import rkck45_coupled_hip_solver as hs

    out = []
    for n, var, ns in cases:
        w = W.serving_bins(seed=21, sizes={n: ns}, periods=periods)[n]
        if n > 4:   # smaller bubbles spread over most of a wavelength
            rng = np.random.default_rng(5)
            w.R0 = rng.uniform(5, 30, size=(ns, n))
            lam = W.wavelength(w.mat, 25.0)
            w.X0 = W.sample_positions(rng, ns, n, 0.0, 0.9, 3 * w.R0.max(1) * 1e-6 / lam)
        m = W.build_model(w, variant=var)
        if event_mode != "bracket":
            m._solver_spec = dataclasses.replace(m._solver_spec, event_mode=event_mode, etol=1e-12)
        m.runtime.host_buffers.state.time_end[:] = 0.05
        m.solve(max_steps=1 << 20, kernel_steps=1 << 20)   # loads the code object
        db = m.runtime.device_buffers
        y0 = db.state.actual_state.clone()
        h0 = db.state.time_step.clone()
        times = []
        for _ in range(reps):
            db.state.actual_state.copy_(y0)
            db.state.time_step.copy_(h0)
            db.state.actual_time.zero_()
            db.state.time_end.fill_(periods)
            db.status.actual_event.fill_(-1.0)
            db.status.total_steps.zero_()
            db.status.status_flags.zero_()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            hs.launch(m.runtime.kernel, db, 1 << 20)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        att = db.status.total_steps.cpu().numpy()
        info = m.runtime.kernel.info()
        ms = min(times)
        out.append({"n": n, "mapping": info.mapping, "event_mode": event_mode, "systems": ns,
                    "kernel_ms": round(ms, 3), "events": int((db.status.status_flags == 2).sum().item()),
                    "max_attempts": int(att.max()), "mean_attempts": round(float(att.mean()), 1),
                    "us_per_attempt": round(ms * 1e3 / max(int(att.max()), 1), 3), "vgprs": info.num_regs,
                    "scratch_bytes": info.local_bytes, "lds_bytes": info.lds_bytes})
    return out


def param_steps(steps: int = 2, periods_per_step: float = 10.0) -> int:
    """Launches of the parameter-study window kernel; returns the attempts made."""
    import torch

    from . import workloads as W
    from .backends.hip.rkck45_coupled_hip_solver import solve_periodic

    w = W.param_study(side=256, periods=200.0, discard=50)
    m = W.build_model(w)
    m.runtime.host_buffers.state.time_end[:] = 0.0
    m._ensure_buffers()
    m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
    db = m.runtime.device_buffers
    max_r = torch.zeros((w.n_windows, w.ns, 2), dtype=torch.float64, device="cuda")
    mean_x = torch.zeros_like(max_r)
    for k in range(steps):
        db.state.time_end.add_(periods_per_step).clamp_(max=200.0)
        solve_periodic(m.runtime.kernel, db, w.period, w.first_window, max_r, mean_x, resume=k > 0)
    torch.cuda.synchronize()
    return int(db.status.total_steps.to(torch.int64).sum().item())


SWEEP_NS = (256, 512, 1024, 2048, 4096, 8192, 16384, 32768)
SWEEP_SEEDS = (11, 42, 55, 77, 89, 119)


def sweep(scene: str = "2B_TYPICAL", ns_list: Sequence[int] = SWEEP_NS, seeds: Sequence[int] = SWEEP_SEEDS,
          kernel_steps: int = 512, measured: int = 5, warmup: int = 1) -> List[Dict]:
    """The reference's NS sweep on one scene; returns one record per NS."""
    import time

    import torch

    from . import workloads as W

    out = []
    for ns in ns_list:
        times, att, periods = [], 0, 0.0
        for seed in seeds:
            w = W.harness_scene(scene, ns, seed)
            m = W.build_model(w)
            m._ensure_buffers()
            db = m.runtime.device_buffers

            def one():
                db.status.actual_event.fill_(-1.0)
                db.status.total_steps.zero_()
                m.solve(max_steps=10 ** 7, kernel_steps=kernel_steps, sync_from_host=True)
            for _ in range(warmup):
                one()
            for _ in range(measured):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                one()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
                att += int(db.status.total_steps.to(torch.int64).sum().item())
                periods += float(db.state.actual_time.sum().item())
        t = float(np.mean(times))
        F = flops_per_attempt(w.n)
        rec = {"scene": scene, "ns": ns, "bubbles": w.n, "kernel_steps": kernel_steps, "seeds": len(seeds),
               "mean_solve_s": t, "std_solve_s": float(np.std(times)), "sim_time_per_wall_s": 5.0 / t,
               "periods_per_s": periods / float(np.sum(times)), "attempts_per_solve": att / len(times),
               "tflops": att * F / float(np.sum(times)) / 1e12,
               "frac": att * F / float(np.sum(times)) / 1e12 / PEAK_FP64_TFLOPS,
               "mapping": m.runtime.kernel.info().mapping}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    return out


def main(argv: Sequence[str] = None) -> None:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("what", choices=["copy", "latency", "param", "sweep"])
    p.add_argument("--scenes", default="2B_TYPICAL", help="sweep: comma list of harness scenes")
    p.add_argument("--mib", type=int, default=512)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--cases", default="", help="latency: comma list like 2lane,4group (default: all)")
    p.add_argument("--systems", default="",
                   help="latency: comma list of system counts run for every case (default: each case's own)")
    p.add_argument("--event-mode", default="bracket", choices=["bracket", "exact"],
                   help="latency: collision-event mode of the solver")
    a = p.parse_args(argv)
    if a.what == "copy":
        print(f"copy kernel: {copy_kernel(a.mib)} bytes read and written per launch", flush=True)
    elif a.what == "param":
        print(f"parameter study: {param_steps(a.steps)} attempts in {a.steps} launches", flush=True)
    elif a.what == "sweep":
        for sc in a.scenes.split(","):
            sweep(sc)
    else:
        cases = LATENCY_CASES
        if a.cases:
            want = set(a.cases.split(","))
            cases = [c for c in LATENCY_CASES if f"{c[0]}{c[1]}" in want]
        if a.systems:
            cases = [(n, var, int(ns)) for n, var, _ in cases for ns in a.systems.split(",")]
        for r in attempt_latency(cases, event_mode=a.event_mode):
            print(json.dumps(r), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, str(__import__("pathlib").Path(__file__).resolve().parents[1]))
    main()
