"""Benchmark: integrated acoustic periods per second on one MI355X.

Workload (default): the parameter study -- 65 536 systems of N=2 coupled
bubbles on a 256 x 256 (PA1, PA2) grid over [0, 2] bar, R0 = (44, 60) um,
water at 25 degC, dual-frequency standing wave (25 + 50 kHz), FP64 Cash--Karp
at rtol = atol = 1e-10, 200 acoustic periods with per-period max R/R0 and
mean x recorded for periods 51..200.

A step integrates every system through ``--periods-per-step`` more periods
(one launch of the period-window kernel; state, t and h stay resident in HBM
between steps).  ``--steps 20 --periods-per-step 10`` is the whole 200-period
job.  Warm-up steps run on a separate copy of the ensemble and are discarded.

value = sum over systems of integrated periods / wall seconds of the timed
steps.  The JSON line also carries the FP64 roofline of the kernel
(algorithmic FLOPs per attempt times attempts, over the kernel time measured
with HIP events on the launch stream) and a CPU baseline (the C oracle, the
same algorithm, OpenMP over systems) timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "coupled-bubble-control_amd"))
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) peak
PEAK_HBM_GBS = 8000.0


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


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--periods-per-step", type=float, default=10.0)
    p.add_argument("--side", type=int, default=256, help="PA grid side (systems = side^2)")
    p.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    p.add_argument("--cpu-systems", type=int, default=2048)
    p.add_argument("--cpu-periods", type=float, default=20.0)
    return p.parse_args()


def log(msg: str) -> None:
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    log("importing torch")
    import numpy as np
    import torch

    from coupledbubble_control import workloads as W
    from coupledbubble_control.backends.hip.rkck45_coupled_hip_solver import solve_periodic

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    P = float(args.periods_per_step)
    total_periods = 200.0
    w = W.param_study(side=args.side, periods=total_periods, discard=50)
    ns = w.ns
    F = flops_per_attempt(w.n)

    def make():
        m = W.build_model(w)
        hb = m.runtime.host_buffers
        hb.state.time_end[:] = 0.0
        m._ensure_buffers()
        m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
        max_r = torch.zeros((w.n_windows, ns, 2), dtype=torch.float64, device=dev)
        mean_x = torch.zeros_like(max_r)
        return m, max_r, mean_x

    def step(m, max_r, mean_x):
        db = m.runtime.device_buffers
        db.state.time_end.add_(P).clamp_(max=total_periods)
        solve_periodic(m.runtime.kernel, db, w.period, w.first_window, max_r, mean_x)

    # warm-up on a discarded copy
    log(f"building {ns} systems")
    mw, rw, xw = make()
    for _ in range(args.warmup):
        step(mw, rw, xw)
    torch.cuda.synchronize()
    del mw, rw, xw

    log("warm-up done")
    m, max_r, mean_x = make()
    db = m.runtime.device_buffers
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_before = db.state.actual_time.clone()
    att_before = db.status.total_steps.to(torch.int64).sum().item()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        step(m, max_r, mean_x)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {wall:.3f} s")
    advanced = (db.state.actual_time - t_before).sum().item()
    attempts = db.status.total_steps.to(torch.int64).sum().item() - att_before
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    flags = db.status.status_flags.cpu().numpy()
    value = advanced / wall
    achieved_tflops = attempts * F / (sum(kernel_ms) * 1e-3) / 1e12
    info = m.runtime.kernel.info()

    cpu = None
    if args.cpu_baseline:
        from oracle import ckm1d_oracle as O

        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = w.subset(np.linspace(0, ns - 1, args.cpu_systems).astype(int))
        sub.t_end[:] = args.cpu_periods
        ms = W.build_model(sub)
        bufs = O.host_buffers_dict(ms)
        spec = O.spec_from_model(ms)
        log(f"cpu baseline: {args.cpu_systems} systems x {args.cpu_periods:g} periods on {nthreads} threads")
        c0 = time.perf_counter()
        out = O.solve_periodic(spec, bufs, w.period, 0, 0)
        cwall = time.perf_counter() - c0
        log(f"cpu baseline done in {cwall:.1f} s")
        cpu = {"value": float(np.sum(bufs["actual_time"]) / cwall), "unit": "periods/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{args.cpu_systems} grid points x {args.cpu_periods:g} periods (same physics, "
                         f"rtol=atol=1e-10), C oracle with OpenMP, {cwall:.1f} s",
               "attempts": int(out["total_steps"].sum())}

    line = {
        "metric": "acoustic periods/s (env-steps/s) on 1 GPU; p50/p99 step latency vs plain",
        "value": value,
        "unit": "periods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "systems": ns, "bubbles_per_system": w.n,
                   "periods_per_step": P, "periods_total": min(total_periods, P * args.steps),
                   "rtol": w.rtol, "atol": w.atol, "field": w.field, "drive_khz": [25.0, 50.0],
                   "window_stats": "max R/R0 and mean x per period, periods 51..200",
                   "mapping": info.mapping, "block_size": info.block_size, "vgprs": info.num_regs,
                   "scratch_bytes": info.local_bytes},
        "roofline": {"bound": "mfma", "pipe": "fp64-valu", "achieved": achieved_tflops,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                     "traffic": None, "flops_per_attempt": F, "attempts": int(attempts),
                     "kernel_ms_mean": float(np.mean(kernel_ms))},
        "cpu_baseline": cpu,
        "status_counts": {str(k): int(v) for k, v in zip(*np.unique(flags, return_counts=True))},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
