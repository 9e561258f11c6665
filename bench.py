"""Benchmark: integrated acoustic periods per second on one MI355X.

Workloads (``--workload``; water at 25 degC, FP64 Cash--Karp, Keller--Miksis
bubbles with implicit radial/translational coupling):

* ``serving`` (default) -- the env-step serving loop: 16 384 resident systems
  (12 288 of N=2, 4 096 of N=4, 25 + 50 kHz standing wave, rtol = atol = 1e-9).
  Requests of 64..1024 envs (one vector env's step: new PA1, PA2 ~ U[0, 1.5] bar,
  integrate K = 5 periods from the resident state and step size) arrive Poisson
  at 200 000 env-steps/s; failed envs and 0.1 % of envs per step (10 % per 100
  steps) restart with fresh radii and positions.  A step is one round of the
  offered load: one request per vector env on average (33 requests, 16 384 env
  steps, ~82 ms of arrivals).  The request stream holds ``--warmup`` rounds,
  then the ``--steps`` timed rounds, then 3 more rounds that keep the load on
  while the timed requests finish; value and latency cover the timed requests
  only (steady state).  Steps are ordered per env, the data dependency of a
  resident system: an env's step of a request starts as soon as its own step of
  the vector env's previous request has finished (the same requests are also
  served with each request waiting for the vector env's whole previous step,
  ``continuous_venv_order``).  The same requests are then served by the
  plain path (one kernel launch per batch, per bubble count), and both paths
  are also driven closed-loop (each vector env resubmits as soon as its step
  completes) for ``--saturation-s`` seconds to measure capacity, and the
  continuous path once more with asynchronous envs (each env resubmits as soon
  as its own step completes, ``saturation.async_*``).  "Plain" is
  static batching: everything queued when the previous batch has finished is
  one batch (one launch per bubble count), and the next batch starts when all
  of its launches have ended; ``plain_per_bin`` also reports one such batch
  loop per bubble count.
  value = integrated periods of the timed requests / the span from their first
  arrival to their last completion, continuous path under the Poisson load;
  latency = request arrival to completion, p50/p99 over the timed requests.
* ``plain`` -- the CPU-runnable case: 256 systems of N=2, one 5-period RL step at
  rtol = atol = 1e-9; a step is one solve of the ensemble from its initial state.
* ``param_study`` -- 65 536 systems of N=2 on a 256 x 256 (PA1, PA2) grid over
  [0, 2] bar, R0 = (44, 60) um, 200 periods with per-period max R/R0 and mean x
  for periods 51..200, rtol = atol = 1e-10; a step is ``--periods-per-step``
  more periods of every system (one launch of the window kernel), by default the
  whole 200-period job in one step, as the reference integrates the horizon in
  one solver call (a launch lasts as long as its slowest system, so 20 steps of
  10 periods add the per-launch tails: 3.09e6 against 3.43e6 periods/s).
* ``mixed`` -- 32 768 systems with N in {2, 3, 4, 8} at 8:4:2:1 and horizons of
  1..50 periods, rtol = atol = 1e-9: the autotuned persistent work-queue
  ensemble (all bins in one fused persistent kernel when the autotuner maps every
  bin to lane groups; the per-bin launches on side streams are timed beside it)
  against one plain launch loop per bubble count; a step is the whole ensemble.
* ``large`` -- 4 096 systems of N=16 (32 x 32 coupling solve per RHS) mixed 1:1
  with 4 096 systems of N=2, 20 periods, rtol = atol = 1e-9; as ``mixed``.
* ``env`` -- the vector environment itself: PosNBC1D with 16 384 envs of 2
  bubbles (``--env-bubbles``, ``--env-count``) at its default physics (R0 = 60 um,
  25 + 50 kHz SW_N drive, 5 periods per step, 128-step episodes, random
  scenes), random actions; a step is one env.step (set_action, solve, observe,
  termination, rewards, advance_time, resets of finished episodes).  Reports
  step latency and the split between the solver and the rest of the step.
* ``rollout`` -- PPO-style rollout collection over PosNBC1D's full step (rewards,
  terminations, time-outs, episode resets): 16 384 envs of 2 bubbles
  (``--env-count``, ``--env-bubbles``), 5 periods per step, actions from an actor
  MLP (256x256 tanh, Beta mean) on the observations.  The synchronous loop (the
  policy on every env, then PosNBC1D.step on every env, as rl/ppo.py:249-300
  does; value: its periods/s) beside ServedPosNBC1D with ``--rollout-groups``
  blocks of envs stepping on their own over the serving kernel and with one
  block; a step = one rollout step of every env.

Every line carries the FP64 roofline of the dominant kernel (algorithmic FLOPs
per attempt -- the count in ``flops_per_attempt`` -- times the attempts the
device counters recorded, over the kernel time) and a CPU baseline: the C
oracle (same algorithm, OpenMP over systems) timed on a bounded sample of the
same workload.

The default serving run then measures plain, param_study, mixed, large and env, each
in a child process at its own defaults without a CPU baseline (``--also``; about a
minute more), and adds their value, unit, ms per step, FP64 roofline fraction and
vs-plain ratio to the line under ``other_workloads``.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "coupled-bubble-control_amd"))
sys.path.insert(0, str(ROOT))

from coupledbubble_control.profiling import PEAK_FP64_TFLOPS, flops_per_attempt  # noqa: E402

# HBM bytes per system of one launch of the parameter-study kernel over the whole 200-period
# job, measured with rocprofv3 (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors of a copy
# kernel of known size): 539 B read + 3 416 B written, the writes mostly the per-period window
# statistics (profiles/lane_kernel_param_study_65536xN2.md).  Only that launch length was
# measured; other --periods-per-step report no traffic.
PARAM_STUDY_TRAFFIC_BYTES_PER_SYSTEM_200 = 3955
# Fabric bytes per env step served by the two persistent serving kernels, measured with
# rocprofv3 over a default serving run (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors of
# a copy kernel of known size), total bytes of the serving dispatches of the per-env-ordered
# continuous run / its env steps; ring, control words and outputs live in mapped host
# memory, so this bounds the HBM bytes from above: 1.28 GB read + 0.27 GB written over
# 458 752 env steps (`--steps 20 --warmup 5 --serving-plain 0`), i.e. 2 784 + 584 bytes
# (profiles/serving_16384_N2x12288_N4x4096.md).
SERVING_TRAFFIC_BYTES_PER_ENV_STEP = 3368
METRIC = "acoustic periods/s (env-steps/s) on 1 GPU; p50/p99 step latency vs plain"


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=["serving", "plain", "param_study", "mixed", "large", "env", "rollout"],
                   default="serving")
    p.add_argument("--steps", type=int, default=None,
                   help="serving: timed rounds of one request per vector env (default 34: 1 122 requests, "
                        "~2.8 s of arrivals); param_study: steps of "
                        "--periods-per-step (default: as many as cover 200 periods); mixed/large: ensemble "
                        "solves (default 1)")
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--rate", type=float, default=200000.0, help="serving: offered env-steps/s")
    p.add_argument("--saturation-s", type=float, default=3.0, help="serving: closed-loop seconds per path (0: off)")
    p.add_argument("--poll-interval", type=int, default=0,
                   help="serving: attempts between ring polls of a busy wave (0: the ServeOptions default)")
    p.add_argument("--serving-plain", type=int, default=1,
                   help="serving: also run the plain paths and the closed loops (0: the continuous path only, "
                        "e.g. under a profiler)")
    p.add_argument("--periods-per-step", type=float, default=200.0,
                   help="param_study: periods integrated per step (one window-kernel launch)")
    p.add_argument("--side", type=int, default=256, help="param_study: PA grid side (systems = side^2)")
    p.add_argument("--env-bubbles", type=int, default=2, help="env: bubbles per env")
    p.add_argument("--env-count", type=int, default=16384, help="env: number of envs")
    p.add_argument("--rollout-groups", type=int, default=4, help="rollout: blocks of envs stepping on their own")
    p.add_argument("--env-trajectories", type=int, default=0,
                   help="env: record trajectories (dense output, device arena, zarr files in a temporary "
                        "directory); 2 writes the files on a background thread")
    p.add_argument("--env-episode-length", type=int, default=128, help="env: steps per episode")
    p.add_argument("--env-policy", type=int, default=0,
                   help="env: choose actions with an actor-critic MLP forward (256x256 tanh, Beta mean) and "
                        "evaluate the critic on the next observations each step, as a PPO rollout does")
    p.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU sample length")
    p.add_argument("--also", default=None,
                   help="serving: comma list of further workloads measured after it in the same run, each "
                        "at its defaults without a CPU baseline, summarised under \"other_workloads\" "
                        "(default plain,param_study,mixed,large,env; '' for none)")
    return p.parse_args()


def log(msg: str) -> None:
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def host_info() -> dict:
    """CPU model and core counts of the host the CPU baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": usable, "os_cpu_count": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def roofline(flops: float, seconds: float, **extra) -> dict:
    achieved = flops / seconds / 1e12 if seconds > 0 else 0.0
    out = {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
           "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None}
    out.update(extra)
    return out


def attempts_of(models) -> dict:
    import torch

    return {m._execution_spec.ups: int(m.runtime.device_buffers.status.total_steps.to(torch.int64).sum().item())
            for m in models}


def cpu_step_sample(W, ws, target_s: float, label: str, threads: int = 0):
    """C oracle (OpenMP) on growing subsets of the given workloads (the whole set at most,
    then repeated) until one pass takes at least target_s / 4; returns the baseline dict.
    threads > 0 runs the oracle's loops on that many threads (1: one core)."""
    import numpy as np

    from oracle import ckm1d_oracle as O

    nthreads = O.set_threads(threads if threads > 0 else cpu_threads())
    frac, reps = 1.0 / 512, 1
    while True:
        periods, wall, systems = 0.0, 0.0, 0
        for _ in range(reps):
            for w in ws:
                k = max(8, int(w.ns * frac))
                sub = w.subset(np.linspace(0, w.ns - 1, k).astype(int))
                m = W.build_model(sub)
                bufs = O.host_buffers_dict(m)
                spec = O.spec_from_model(m)
                c0 = time.perf_counter()
                O.solve(spec, bufs, kernel_steps=1 << 20, max_steps=1 << 22)
                wall += time.perf_counter() - c0
                periods += float(np.sum(bufs["actual_time"]))
                systems += k
        if wall >= target_s / 4:
            break
        grow = min(8.0, max(2.0, target_s / max(wall, 1e-3)))
        if frac < 1.0:
            frac = min(1.0, frac * grow)
        else:
            reps = int(np.ceil(reps * grow))
    return {"value": periods / wall, "unit": "periods/s", "cores": nthreads, "kind": "port",
            "sample": f"{systems} system runs of {label} ({periods:.0f} periods), C oracle "
                      f"{'with OpenMP' if nthreads > 1 else 'on one core'}, {wall:.1f} s"}


# ---------------------------------------------------------------------------
def bench_serving(args):
    import numpy as np
    import torch

    from coupledbubble_control import workloads as W
    from coupledbubble_control.backends.hip.serve import ServeOptions
    from coupledbubble_control.serving import ServingEngine, c0_scene, closed_loop, make_requests, vector_envs

    sizes = {2: 12288, 4: 4096}
    K = 5.0
    venvs = vector_envs(sizes, np.random.default_rng(3))
    V = len(venvs)
    # A step is one round of the offered load: V requests, on average one env step of
    # every vector env (16 384 env steps).  The request stream holds `warmup` rounds that
    # load the queues, then the timed rounds, then COOL rounds that keep the load on
    # while the timed requests finish; only the timed requests are measured.
    steps = args.steps or 34
    COOL = 3
    n_req = (args.warmup + steps + COOL) * V
    lo, hi = args.warmup * V, (args.warmup + steps) * V
    lam = W.wavelength(W.water_25c(), 25.0)

    def engine(mode, per_bin=False, order="env"):
        bins_w = W.serving_bins(seed=3, sizes=sizes, periods=K)
        bins = {n: W.build_model(w) for n, w in bins_w.items()}
        eng = ServingEngine(bins, periods=K, mode=mode, reset_prob=0.001,
                            reset_scene=lambda n, k, rng: c0_scene(n, k, rng, lam),
                            serve_opts=sopts, seed=3, plain_per_bin=per_bin, order=order)
        return eng, bins

    sopts = ServeOptions(max_attempts=1 << 20, **({"poll_interval": args.poll_interval} if args.poll_interval else {}))

    def run(mode, reqs=None, closed=None, per_bin=False, window=True, order="env", asynchronous=False):
        eng, bins = engine(mode, per_bin, order)
        a0 = attempts_of(bins.values())
        gc.collect()
        gc.disable()   # no collector pass inside the timed request loop (as timeit does)
        try:
            if asynchronous:
                st = eng.run_async(closed, seed=9, timeout_s=300.0, progress=log)
            elif closed is not None:
                first, feeder = closed_loop(venvs, closed, seed=9)
                st = eng.run(first, timeout_s=300.0, progress=log, feeder=feeder)
            else:
                st = eng.run(reqs, timeout_s=300.0, progress=log)
            win = eng.window(lo, hi) if window and closed is None else None
        finally:
            gc.enable()
            eng.close()
        a1 = attempts_of(bins.values())
        st["attempts"] = {n: a1[n] - a0[n] for n in a1}
        st["flops"] = float(sum(st["attempts"][n] * flops_per_attempt(n) for n in a1))
        info = {n: m.runtime.kernel.info() for n, m in bins.items()}
        return st, win, info

    log(f"serving: {sum(sizes.values())} resident systems, {V} vector envs; a step = {V} requests; "
        f"{args.warmup} + {steps} + {COOL} rounds")
    run("continuous", make_requests(venvs, args.rate, count=V, seed=100), window=False)   # module load, first launch
    torch.cuda.synchronize()
    reqs = lambda: make_requests(venvs, args.rate, count=n_req, seed=3)   # noqa: E731
    log(f"continuous: {n_req} requests at {args.rate:g} env-steps/s")
    whole, cont, info = run("continuous", reqs())
    log(f"continuous window: {cont['periods_per_s']:.4g} periods/s p50 {cont['p50_ms']:.2f} ms "
        f"p99 {cont['p99_ms']:.2f} ms over {cont['requests']} requests")
    log("continuous, steps ordered per vector env: same requests")
    _, cont_v, _ = run("continuous", reqs(), order="venv")
    log(f"continuous (vector-env order) window: {cont_v['periods_per_s']:.4g} periods/s "
        f"p50 {cont_v['p50_ms']:.2f} ms p99 {cont_v['p99_ms']:.2f} ms")
    plain = plain_bin = sat = None
    if args.serving_plain:
        log("plain (static batches): same requests")
        _, plain, _ = run("plain", reqs())
        log(f"plain window: {plain['periods_per_s']:.4g} periods/s p50 {plain['p50_ms']:.2f} ms "
            f"p99 {plain['p99_ms']:.2f} ms")
        log("plain, one batch loop per bubble count: same requests")
        _, plain_bin, _ = run("plain", reqs(), per_bin=True)
        log(f"plain per bin window: {plain_bin['periods_per_s']:.4g} periods/s p50 {plain_bin['p50_ms']:.2f} ms")
        if args.saturation_s > 0:
            log(f"closed loop {args.saturation_s:g} s: continuous")
            sc, _, _ = run("continuous", closed=args.saturation_s)
            log(f"closed loop {args.saturation_s:g} s: plain")
            sp, _, _ = run("plain", closed=args.saturation_s)
            log(f"closed loop {args.saturation_s:g} s: continuous, asynchronous envs")
            sa, _, _ = run("continuous", closed=args.saturation_s, asynchronous=True)
            sat = {"continuous_periods_per_s": sc["periods_per_s"], "plain_periods_per_s": sp["periods_per_s"],
                   "ratio": sc["periods_per_s"] / sp["periods_per_s"],
                   "continuous_p50_ms": sc["p50_ms"], "continuous_p99_ms": sc["p99_ms"],
                   "plain_p50_ms": sp["p50_ms"], "plain_p99_ms": sp["p99_ms"],
                   "continuous_tflops": sc["flops"] / sc["wall_s"] / 1e12,
                   "continuous_frac": sc["flops"] / sc["wall_s"] / 1e12 / PEAK_FP64_TFLOPS,
                   "async_periods_per_s": sa["periods_per_s"],
                   "async_ratio": sa["periods_per_s"] / sp["periods_per_s"],
                   "async_p50_ms": sa["p50_ms"], "async_p99_ms": sa["p99_ms"],
                   "async_env_steps": sa["env_steps"],
                   "async_tflops": sa["flops"] / sa["wall_s"] / 1e12,
                   "async_frac": sa["flops"] / sa["wall_s"] / 1e12 / PEAK_FP64_TFLOPS,
                   "note": "closed loop: each vector env submits its next step when its step completes "
                           "(ratio: continuous / plain); asynchronous: each env submits its next step "
                           "when its own step completes (async_ratio against the same plain loop, whose "
                           "static batches hold every queued env either way)"}

    cpu = cpu1 = None
    if args.cpu_baseline:
        ws = list(W.serving_bins(seed=3, sizes=sizes, periods=K).values())
        log("cpu baseline (all cores)")
        cpu = cpu_step_sample(W, ws, args.cpu_seconds, "the resident set, one 5-period step each")
        log("cpu baseline (one core)")
        cpu1 = cpu_step_sample(W, ws, args.cpu_seconds * 0.6, "the resident set, one 5-period step each",
                               threads=1)
        cpu = with_one_core(cpu, cpu1)
    keys = ("wall_s", "requests", "env_steps", "periods", "periods_per_s", "env_steps_per_s", "p50_ms", "p99_ms",
            "mean_ms", "max_ms", "failed_env_steps")
    att = cont["attempts"]
    flops = float(sum(att[n] * flops_per_attempt(n) for n in att))
    traffic = (SERVING_TRAFFIC_BYTES_PER_ENV_STEP * cont["env_steps"]
               if args.rate == 200000.0 and SERVING_TRAFFIC_BYTES_PER_ENV_STEP else None)
    return {
        "value": cont["periods_per_s"],
        "unit": "periods/s",
        "steps": steps,
        "ms_per_step": cont["wall_s"] * 1e3 / steps,
        "config": {"workload": "serving_16384_N2x12288_N4x4096", "resident_systems": 16384,
                   "bins": {"2": 12288, "4": 4096}, "vector_envs": V, "burst_envs": [64, 1024],
                   "offered_env_steps_per_s": args.rate, "periods_per_step": K, "rtol": 1e-9, "atol": 1e-9,
                   "field": "SW_N", "drive_khz": [25.0, 50.0], "reset": "failed envs + 10% per 100 steps",
                   "step": f"one round of {V} requests (every vector env once on average, 16 384 env steps)",
                   "window": f"requests {lo}..{hi - 1} of {n_req}: after {args.warmup} warm-up rounds, "
                             f"followed by {COOL} rounds that keep the load on",
                   "mapping": {str(n): i.mapping for n, i in info.items()},
                   "vgprs": {str(n): i.num_regs for n, i in info.items()},
                   "poll_interval": sopts.poll_interval, "idle_us": sopts.idle_us,
                   "order": "per env: an env's step of a request starts when its own step of the vector "
                            "env's previous request has finished (continuous_venv_order: when the whole "
                            "previous request has)"},
        "latency_ms": {"p50": cont["p50_ms"], "p99": cont["p99_ms"],
                       "samples": cont["requests"],
                       **({"plain_p50": plain["p50_ms"], "plain_p99": plain["p99_ms"]} if plain else {})},
        "vs_plain": cont["periods_per_s"] / plain["periods_per_s"] if plain else None,
        "env_steps_per_s": cont["env_steps_per_s"],
        "continuous": {k: cont[k] for k in keys},
        "continuous_venv_order": {k: cont_v[k] for k in keys},
        "plain": {k: plain[k] for k in keys} if plain else None,
        "plain_per_bin": {k: plain_bin[k] for k in keys} if plain_bin else None,
        "saturation": sat,
        "whole_run": {"requests": whole["requests"], "env_steps": whole["env_steps"], "wall_s": whole["wall_s"],
                      "kernel_launches": whole.get("kernel_launches"),
                      "attempts": {str(n): v for n, v in whole["attempts"].items()}},
        "roofline": roofline(flops, cont["wall_s"], window="the timed requests' span (persistent serving kernels)",
                             flops_per_attempt={str(n): flops_per_attempt(n) for n in sizes},
                             attempts={str(n): v for n, v in att.items()},
                             traffic=traffic,
                             traffic_note="rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 of both serving kernels "
                                          "over a default run, per env step served, times the window's env steps; "
                                          "mostly ring polls of mapped host memory (an upper bound on HBM bytes)"),
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
def bench_plain(args):
    """The CPU-runnable plain workload: 256 systems of N = 2 (R0 ~ U[20, 80] um, PA ~ U[0, 1.5]
    bar, 25 + 50 kHz standing wave), one RL step of 5 periods at rtol = atol = 1e-9; a step
    is one solve of the whole ensemble from its initial state (model.solve, kernel_steps 2048)."""
    import numpy as np
    import torch

    from coupledbubble_control import workloads as W

    steps = args.steps or 20
    w = W.plain_cpu(seed=0)
    m = W.build_model(w)
    m._ensure_buffers()
    db = m.runtime.device_buffers

    def one():
        db.status.actual_event.fill_(-1.0)
        m.solve(max_steps=10 ** 7, kernel_steps=2048, sync_from_host=True)

    for _ in range(args.warmup):
        one()
    torch.cuda.synchronize()
    db.status.total_steps.zero_()
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    wall = float(np.sum(times))
    att = int(db.status.total_steps.to(torch.int64).sum().item())
    periods = float(w.ns * 5.0 * steps)
    cpu = None
    if args.cpu_baseline:
        cpu = with_one_core(cpu_step_sample(W, [w], args.cpu_seconds, "the 256-system plain workload"),
                            cpu_step_sample(W, [w], args.cpu_seconds * 0.5, "the 256-system plain workload",
                                            threads=1))
    F = flops_per_attempt(2)
    return {
        "value": periods / wall, "unit": "periods/s", "steps": steps, "ms_per_step": wall * 1e3 / steps,
        "config": {"workload": "plain_cpu_256xN2: 256 systems x N=2, 5 periods, rtol=atol=1e-9", "systems": w.ns,
                   "mapping": m.runtime.kernel.info().mapping},
        "roofline": roofline(att * F, wall, window="timed solves (host launch loop included)",
                             flops_per_attempt=F, attempts=att),
        "cpu_baseline": cpu,
    }


def with_one_core(cpu: dict, one: dict) -> dict:
    """The all-core baseline with its one-core row and the host it ran on."""
    cpu["one_core"] = {k: one[k] for k in ("value", "cores", "sample")}
    cpu["host"] = host_info()
    return cpu


def _param_cpu_sample(w, ns: int, target_s: float, threads: int = 0) -> dict:
    """The C oracle's window integration on growing subsets of the parameter-study grid
    (20 periods each) until one run takes at least target_s / 3."""
    import numpy as np

    from coupledbubble_control import workloads as W
    from oracle import ckm1d_oracle as O

    nthreads = O.set_threads(threads if threads > 0 else cpu_threads())
    cpu_systems, cpu_periods = (2048 if nthreads > 1 else 128), 20.0
    for _ in range(6):
        sub = w.subset(np.linspace(0, ns - 1, cpu_systems).astype(int))
        sub.t_end[:] = cpu_periods
        ms = W.build_model(sub)
        bufs = O.host_buffers_dict(ms)
        spec = O.spec_from_model(ms)
        log(f"cpu baseline: {cpu_systems} systems x {cpu_periods:g} periods on {nthreads} threads")
        c0 = time.perf_counter()
        O.solve_periodic(spec, bufs, w.period, 0, 0)
        cwall = time.perf_counter() - c0
        if cwall >= target_s / 3 or cpu_systems >= ns:
            break
        cpu_systems = min(ns, int(cpu_systems * min(8.0, target_s / max(cwall, 1e-3))))
    return {"value": float(np.sum(bufs["actual_time"]) / cwall), "unit": "periods/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{cpu_systems} grid points x {cpu_periods:g} periods (same physics, rtol=atol=1e-10), "
                      f"C oracle {'with OpenMP' if nthreads > 1 else 'on one core'}, {cwall:.1f} s"}


def bench_param_study(args):
    import numpy as np
    import torch

    from coupledbubble_control import workloads as W
    from coupledbubble_control.backends.hip.rkck45_coupled_hip_solver import solve_periodic

    dev = torch.device("cuda:0")
    P = float(args.periods_per_step)
    total_periods = 200.0
    steps = args.steps or max(1, int(round(total_periods / P)))
    w = W.param_study(side=args.side, periods=total_periods, discard=50)
    ns = w.ns
    F = flops_per_attempt(w.n)

    def make():
        m = W.build_model(w)
        m.runtime.host_buffers.state.time_end[:] = 0.0
        m._ensure_buffers()
        m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
        max_r = torch.zeros((w.n_windows, ns, 2), dtype=torch.float64, device=dev)
        return m, max_r, torch.zeros_like(max_r)

    def step(m, max_r, mean_x, k):
        # later steps continue one integration: a system stopped by a failure stays stopped
        db = m.runtime.device_buffers
        db.state.time_end.add_(P).clamp_(max=total_periods)
        solve_periodic(m.runtime.kernel, db, w.period, w.first_window, max_r, mean_x, resume=k > 0)

    log(f"param study: building {ns} systems")
    mw, rw, xw = make()
    for k in range(args.warmup):
        step(mw, rw, xw, k)
    torch.cuda.synchronize()
    del mw, rw, xw
    log("warm-up done")
    m, max_r, mean_x = make()
    db = m.runtime.device_buffers
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t_before = db.state.actual_time.clone()
    att_before = db.status.total_steps.to(torch.int64).sum().item()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        step(m, max_r, mean_x, k)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {steps} steps in {wall:.3f} s")
    advanced = (db.state.actual_time - t_before).sum().item()
    attempts = db.status.total_steps.to(torch.int64).sum().item() - att_before
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    flags = db.status.status_flags.cpu().numpy()
    info = m.runtime.kernel.info()

    cpu = None
    if args.cpu_baseline:
        cpu = _param_cpu_sample(w, ns, args.cpu_seconds)
        cpu1 = _param_cpu_sample(w, ns, args.cpu_seconds * 0.5, threads=1)
        cpu = with_one_core(cpu, cpu1)
    return {
        "value": advanced / wall,
        "unit": "periods/s",
        "steps": steps,
        "ms_per_step": wall * 1e3 / steps,
        "config": {"workload": w.name, "systems": ns, "bubbles_per_system": w.n, "periods_per_step": P,
                   "periods_total": min(total_periods, P * steps), "rtol": w.rtol, "atol": w.atol,
                   "field": w.field, "drive_khz": [25.0, 50.0],
                   "window_stats": "max R/R0 and mean x per period, periods 51..200",
                   "mapping": info.mapping, "block_size": info.block_size, "vgprs": info.num_regs,
                   "scratch_bytes": info.local_bytes},
        "roofline": roofline(attempts * F / steps, float(np.mean(kernel_ms)) * 1e-3,
                             window="kernel (HIP events per launch)", flops_per_attempt=F, attempts=int(attempts),
                             kernel_ms_mean=float(np.mean(kernel_ms)),
                             traffic=(PARAM_STUDY_TRAFFIC_BYTES_PER_SYSTEM_200 * ns
                                      if P == total_periods and steps == 1 and ns == 65536 else None),
                             traffic_note="HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE of the "
                                          "200-period launch of 65 536 systems"),
        "cpu_baseline": cpu,
        "status_counts": {str(k): int(v) for k, v in zip(*np.unique(flags, return_counts=True))},
    }


# ---------------------------------------------------------------------------
def bench_ensemble(args, large: bool):
    import numpy as np
    import torch

    from coupledbubble_control import workloads as W
    from coupledbubble_control.ensemble import MixedEnsemble, autotune

    ws = W.large_coupling() if large else W.mixed_uneven()
    name = ("large_coupling_4096xN16_4096xN2" if large
            else "mixed_uneven_32768_" + "_".join(f"N{w.n}x{w.ns}" for w in ws))
    steps = args.steps or 1

    def models():
        return [W.build_model(w) for w in ws]

    log(f"{name}: autotune")
    tune = autotune(models(), horizon=0.5, limit=4096)
    tuning = {n: r.choice for n, r in tune.items()}
    log(f"tuning: { {n: (c.variant, c.refill, c.block or 64) for n, c in tuning.items()} }")

    # warm-up: one short ensemble run
    for _ in range(args.warmup):
        ms = models()
        for m in ms:
            m.runtime.host_buffers.state.time_end[:] = np.minimum(m.runtime.host_buffers.state.time_end, 0.25)
        MixedEnsemble(ms, tuning).solve()
        torch.cuda.synchronize()

    def timed_queued(fused="auto"):
        ms = models()
        ens = MixedEnsemble(ms, tuning, fused=fused)
        ens._prepare()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rounds = ens.solve()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ms, ens.fused_active

    def timed_plain():
        ms = models()
        for m in ms:
            m._ensure_buffers()
            m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for m in ms:
            m.solve(max_steps=1 << 22, kernel_steps=1 << 20)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ms

    walls, flops, periods = [], 0.0, 0.0
    for k in range(steps):
        log(f"queued ensemble solve {k + 1}/{steps}")
        wall, ms, fused = timed_queued()
        walls.append(wall)
        att = attempts_of(ms)
        flops += sum(att[n] * flops_per_attempt(n) for n in att)
        periods += float(sum(m.runtime.device_buffers.state.actual_time.sum().item() for m in ms))
    wall = sum(walls)
    launch = "one fused persistent kernel" if fused else "one persistent launch per bin on side streams"
    log(f"queued ({launch}): {periods / wall:.4g} periods/s")
    side = None
    if fused:
        log("queued, one persistent launch per bin on side streams")
        swall, sms, _ = timed_queued("off")
        side = float(sum(m.runtime.device_buffers.state.actual_time.sum().item() for m in sms)) / swall
        log(f"per-bin launches: {side:.4g} periods/s")
    log("plain: one launch loop per bubble count")
    pwall, pms = timed_plain()
    pperiods = float(sum(m.runtime.device_buffers.state.actual_time.sum().item() for m in pms))
    log(f"plain: {pperiods / pwall:.4g} periods/s")
    flags = {w.n: np.unique(m.runtime.device_buffers.status.status_flags.cpu().numpy(), return_counts=True)
             for w, m in zip(ws, ms)}
    cpu = None
    if args.cpu_baseline:
        log("cpu baseline")
        cpu = with_one_core(cpu_step_sample(W, ws, args.cpu_seconds, name),
                            cpu_step_sample(W, ws, args.cpu_seconds * 0.5, name, threads=1))
    return {
        "value": periods / wall,
        "unit": "periods/s",
        "steps": steps,
        "ms_per_step": wall * 1e3 / steps,
        "config": {"workload": name, "bins": {str(w.n): w.ns for w in ws}, "rtol": 1e-9, "atol": 1e-9,
                   "field": "SW_N", "drive_khz": [25.0, 37.5] if large else [25.0, 50.0],
                   "tuning": {str(n): {"mapping": c.variant, "refill": c.refill, "waves": c.waves,
                                       "block": c.block or 64} for n, c in tuning.items()},
                   "tuning_trials_s": {str(n): r.trials for n, r in tune.items()}},
        "vs_plain": (periods / wall) / (pperiods / pwall),
        "plain_periods_per_s": pperiods / pwall,
        "launch": launch,
        "per_bin_launches_periods_per_s": side,
        "roofline": roofline(flops / steps, wall / steps, window=f"timed region ({launch})",
                             flops_per_attempt={str(w.n): flops_per_attempt(w.n) for w in ws}),
        "cpu_baseline": cpu,
        "status_counts": {str(n): {str(int(a)): int(b) for a, b in zip(*f)} for n, f in flags.items()},
    }


def _env_oracle_sample(n: int, T: float, target_s: float, threads: int = 0) -> dict:
    """The oracle environment (oracle/env_step.py) with the C oracle solver on a
    growing number of envs of the env workload, random actions, until one run of
    three steps takes at least target_s / 4."""
    import numpy as np

    from coupledbubble_control.models import CKM1D
    from oracle import ckm1d_oracle as C
    from oracle import env_step as E

    nthreads = C.set_threads(threads if threads > 0 else cpu_threads())

    k = 16
    rng = np.random.default_rng(1)
    while True:
        m = CKM1D(num_systems=k, systems_per_block=64, units_per_system=n, R0=np.full((k, n), 60.0),
                  PA=np.tile([[1.0], [0.0]], k), FR=np.tile([[25.0], [50.0]], k), PS=np.zeros((2, k)),
                  REF_FREQ=25.0, acoustic_field="SW_N")
        parts = [E.Part("XT", [-0.5] * n, [0.5] * n), E.Part("X", [-0.5] * n, [0.5] * n, stack=2)]
        env = E.EnvOracle(k, n, parts, action_rows=[0, 1, 4, 5], action_scales=[1e5, 1e5, np.pi, np.pi],
                          pa_cols=[0, 1])
        sp = E.SceneParams(num_bubbles=n, domain_min=-0.5, domain_max=0.5)
        bufs = C.host_buffers_dict(m)
        bufs["time_end"][:] = T
        env.reset_envs(np.arange(k), E.random_uniform_positions(sp, rng.random((k, n - 1))),
                       E.random_uniform_positions(sp, rng.random((k, n - 1))), bufs["actual_state"])
        spec = C.spec_from_model(m)
        c0 = time.perf_counter()
        for _ in range(3):
            act = rng.random((k, 4)).astype(np.float32)
            env.set_action(act, bufs["control_params"])
            C.solve(spec, bufs, kernel_steps=2048, max_steps=10 ** 6)
            r = env.post(bufs["actual_state"], bufs["actual_time"], bufs["time_step"], bufs["actual_event"],
                         bufs["status_flags"], act)
            d = r["done_ids"]
            if d.size:
                env.reset_envs(d, E.random_uniform_positions(sp, rng.random((d.size, n - 1))),
                               E.random_uniform_positions(sp, rng.random((d.size, n - 1))), bufs["actual_state"])
        wall = time.perf_counter() - c0
        if wall >= target_s / 4 or k >= 1 << 14:
            break
        k = min(1 << 14, int(k * min(8.0, max(2.0, target_s / max(wall, 1e-3)))))
    return {"value": k * T * 3 / wall, "unit": "periods/s", "cores": nthreads, "kind": "port",
            "sample": f"3 env steps of {k} envs x {n} bubbles (oracle environment, C oracle solver "
                      f"{'with OpenMP' if nthreads > 1 else 'on one core'}), {wall:.1f} s"}


def bench_env(args):
    """PosNBC1D steps: latency per env.step and the solver / pipeline split."""
    import numpy as np
    import torch

    from coupledbubble_control.envs import PosNBC1D

    n, ne = args.env_bubbles, args.env_count
    steps = args.steps or 20
    import shutil
    import tempfile

    traj = bool(args.env_trajectories)
    store = tempfile.mkdtemp(prefix="cbc_traj_") if traj else None
    log(f"env: building PosNBC1D {ne} envs x {n} bubbles" + (f", trajectories into {store}" if traj else ""))
    env = PosNBC1D(num_bubbles=n, num_envs=ne, envs_per_block=64, verbose=False, collect_trajectories=traj,
                   trajectory_buffer_store=store or "PosNBC1D", episode_length=args.env_episode_length,
                   trajectory_async_write=args.env_trajectories == 2)
    T = float(env.time_step_length)
    env.reset()
    stream = torch.cuda.current_stream()
    ev = {"physics": [], "post": [], "reset": [], "append": []}
    write_s = []

    def timed(fn, key):
        def run(*a, **kw):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            r = fn(*a, **kw)
            e1.record(stream)
            ev[key].append((e0, e1))
            return r
        return run

    env.step_physics = timed(env.step_physics, "physics")
    env._kernels.post = timed(env._kernels.post, "post")
    env._kernels.reset = timed(env._kernels.reset, "reset")
    if traj:
        tb = env.trajectory_buffer
        tb.step = timed(tb.step, "append")
        end_episode = tb.end_episode

        def timed_end(*a, **kw):
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            end_episode(*a, **kw)
            write_s.append(time.perf_counter() - c0)
        tb.end_episode = timed_end
    policy = None
    if args.env_policy:
        A, D = env.action_space.space_data.shape[0], env.observation_space.space_data.shape[0]

        def mlp(out):
            return torch.nn.Sequential(torch.nn.Linear(D, 256), torch.nn.Tanh(), torch.nn.Linear(256, 256),
                                       torch.nn.Tanh(), torch.nn.Linear(256, out)).cuda()
        actor, critic = mlp(2 * A), mlp(1)
        obs_now = [env._observations]

        def policy():
            with torch.no_grad():
                ab = torch.nn.functional.softplus(actor(obs_now[0])) + 1.0
                return ab[:, :A] / (ab[:, :A] + ab[:, A:])

        def value(o):
            with torch.no_grad():
                return critic(o)
    ev["policy"] = []

    def step_once():
        if not args.env_policy:
            return env.step(None)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        act = policy()
        e1.record(stream)
        r = env.step(act)
        e2, e3 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e2.record(stream)
        value(r[0])
        e3.record(stream)
        ev["policy"].append((e0, e1))
        ev["policy"].append((e2, e3))
        obs_now[0] = r[0]
        return r

    log(f"env: {args.warmup} warmup steps")
    for _ in range(args.warmup):
        step_once()
    torch.cuda.synchronize()
    for v in ev.values():
        v.clear()
    write_s.clear()
    written0 = env.trajectory_buffer.episode_count if traj else 0
    a0 = attempts_of([env.model])[n]
    lat, dones = [], 0
    log(f"env: {steps} timed steps")
    gc.collect()
    gc.disable()   # no collector pass inside a timed step (as timeit does)
    t0 = time.perf_counter()
    for k in range(steps):
        s0 = time.perf_counter()
        info = step_once()[4]
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - s0)
        dones += int(info["dones"].numel()) if info else 0
    wall = time.perf_counter() - t0
    gc.enable()
    att = attempts_of([env.model])[n] - a0
    flush_s = 0.0
    if traj:
        c0 = time.perf_counter()
        env.trajectory_buffer.flush()
        flush_s = time.perf_counter() - c0

    def ms(key):
        return [e0.elapsed_time(e1) for e0, e1 in ev[key]]

    phys, post, rst, app = ms("physics"), ms("post"), ms("reset"), ms("append")
    lat_ms = np.array(lat) * 1e3
    flops = float(att) * flops_per_attempt(n)
    out = {
        "value": ne * T * steps / wall, "unit": "periods/s", "steps": steps, "ms_per_step": wall * 1e3 / steps,
        "config": {"workload": f"env: PosNBC1D {ne} envs x {n} bubbles, {T:g} periods per step, random actions",
                   "envs": ne, "bubbles": n, "periods_per_step": T},
        "env_steps_per_s": ne * steps / wall,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "split_ms_per_step": {"solver": float(np.mean(phys)), "post_kernel_and_counts": float(np.mean(post)),
                              "reset": float(np.mean(rst)) if rst else 0.0,
                              "host_rest": float(np.mean(lat_ms) - np.mean(phys) - np.mean(post)
                                                 - (np.sum(rst) / steps if rst else 0.0))},
        "episodes_finished": dones,
        "episode_length": args.env_episode_length,
        **({"policy_forward_ms_per_step": float(np.sum(ms("policy")) / steps)} if args.env_policy else {}),
        "roofline": roofline(flops, float(np.sum(phys)) / 1e3, window="solver launches of the timed steps",
                             attempts={str(n): att}),
    }
    if traj:
        written = env.trajectory_buffer.episode_count - written0
        nfiles = sum(len(f) for _, _, f in os.walk(store))
        out["trajectories"] = {
            "append_ms_per_step": float(np.mean(app)), "episodes_written": written,
            "end_episode_s_total": float(np.sum(write_s)),
            "episodes_per_s_written": written / float(np.sum(write_s)) if write_s else None,
            "files": nfiles, "dense_points_per_solve": int(env.model._execution_spec.ndo),
            "background_writer": args.env_trajectories == 2, "flush_after_timed_steps_s": flush_s}
        shutil.rmtree(store, ignore_errors=True)
    if args.cpu_baseline:
        log("env: CPU baseline (oracle environment)")
        out["cpu_baseline"] = with_one_core(_env_oracle_sample(n, T, args.cpu_seconds),
                                            _env_oracle_sample(n, T, args.cpu_seconds * 0.5, threads=1))
    return out


def bench_rollout(args):
    """PPO-style rollout collection over PosNBC1D's full step (rewards, terminations,
    time-outs, episode resets): the synchronous loop (policy on every env, then
    PosNBC1D.step on every env, as rl/ppo.py:249-300) against ServedPosNBC1D with
    --rollout-groups blocks stepping on their own over the serving kernel."""
    import numpy as np
    import torch

    from coupledbubble_control.envs import PosNBC1D, ServedPosNBC1D

    n = args.env_bubbles
    ne = args.env_count
    steps = args.steps or 20
    G = args.rollout_groups
    torch.manual_seed(0)

    def make_env():
        return PosNBC1D(num_bubbles=n, num_envs=ne, envs_per_block=64, verbose=False, collect_trajectories=False,
                        episode_length=args.env_episode_length)

    probe = make_env()
    A, D = probe.action_space.space_data.shape[0], probe.observation_space.space_data.shape[0]
    T = float(probe.time_step_length)
    del probe

    def mlp(out):
        return torch.nn.Sequential(torch.nn.Linear(D, 256), torch.nn.Tanh(), torch.nn.Linear(256, 256),
                                   torch.nn.Tanh(), torch.nn.Linear(256, out)).cuda()
    actor = mlp(2 * A)
    calls = [0]

    def policy(obs, begin=0):
        calls[0] += 1
        with torch.no_grad():
            ab = torch.nn.functional.softplus(actor(obs)) + 1.0
            return ab[:, :A] / (ab[:, :A] + ab[:, A:])

    def synchronous():
        env = make_env()
        obs, _ = env.reset()
        for _ in range(max(args.warmup, 1)):
            obs = env.step(policy(obs))[0]
        a0 = attempts_of([env.model])[n]
        calls[0] = 0
        torch.cuda.synchronize()
        gc.collect()
        gc.disable()
        t0 = time.perf_counter()
        for _ in range(steps):
            obs = env.step(policy(obs))[0]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gc.enable()
        return {"wall_s": wall, "attempts": attempts_of([env.model])[n] - a0, "policy_calls": calls[0]}

    def served(groups, lockstep=False):
        env = make_env()
        env.reset()
        sv = ServedPosNBC1D(env, groups=groups)
        try:
            sv.rollout(policy, max(args.warmup, 1), lockstep=lockstep, record=False)
            a0 = attempts_of([env.model])[n]
            calls[0] = 0
            gc.collect()
            gc.disable()
            r = sv.rollout(policy, steps, lockstep=lockstep, record=False)
            gc.enable()
            att = attempts_of([env.model])[n] - a0
        finally:
            sv.close()
        return {"wall_s": r["wall_s"], "attempts": att, "policy_calls": calls[0],
                "group_done_s": [round(float(x), 4) for x in r["group_done_s"]]}

    log(f"rollout: {ne} envs x {n} bubbles x {steps} steps, synchronous PosNBC1D.step loop")
    sync = synchronous()
    log(f"rollout: served, {G} blocks on their own")
    grp = served(G)
    log("rollout: served, one block")
    one = served(1)
    periods = ne * steps * T
    cpu = None
    if args.cpu_baseline:
        log("env: CPU baseline (oracle environment)")
        cpu = with_one_core(_env_oracle_sample(n, T, args.cpu_seconds),
                            _env_oracle_sample(n, T, args.cpu_seconds * 0.5, threads=1))
    F = flops_per_attempt(n)
    best = min((sync, "synchronous"), (grp, f"served_{G}_blocks"), (one, "served_1_block"),
               key=lambda x: x[0]["wall_s"])
    return {
        "value": periods / sync["wall_s"],
        "unit": "periods/s",
        "steps": steps,
        "ms_per_step": sync["wall_s"] * 1e3 / steps,
        "config": {"workload": f"rollout: PosNBC1D {ne} envs x {n} bubbles, {T:g} periods per step, actor MLP "
                               f"policy; the synchronous loop, beside {G} served blocks and one",
                   "envs": ne, "bubbles": n, "periods_per_step": T,
                   "groups": G, "policy": "actor MLP 256x256 tanh, Beta mean, on the observations",
                   "episode_length": args.env_episode_length},
        "env_steps_per_s": ne * steps / sync["wall_s"],
        "synchronous": {**sync, "periods_per_s": periods / sync["wall_s"], "ms_per_step": sync["wall_s"] * 1e3 / steps},
        "served_blocks": {**grp, "periods_per_s": periods / grp["wall_s"], "ms_per_step": grp["wall_s"] * 1e3 / steps},
        "served_one_block": {**one, "periods_per_s": periods / one["wall_s"],
                             "ms_per_step": one["wall_s"] * 1e3 / steps},
        "vs_synchronous": sync["wall_s"] / grp["wall_s"],
        "fastest": best[1],
        "roofline": roofline(sync["attempts"] * F, sync["wall_s"], window="timed synchronous rollout",
                             flops_per_attempt=F, attempts=int(sync["attempts"])),
        "cpu_baseline": cpu,
    }


OTHER_DEFAULT = "plain,param_study,mixed,large,env"


def other_workloads(args, names) -> dict:
    """Each named workload at its own defaults (no CPU baseline, at most one warm-up step),
    summarised: value, unit, ms per step, steps, FP64 roofline fraction and the vs-plain
    ratio where it has one.  Each runs in a child process of its own, as it would run
    alone: the HIP runtime maps streams to the device's four hardware queues by how many
    streams the process created before, and the mixed and large ensembles put their bins
    on side streams that must land on separate queues (after the serving runs' streams,
    two bins of the mixed ensemble shared one and its step took 8.3 s instead of 6.2 s)."""
    import subprocess

    known = ("plain", "param_study", "env", "rollout", "mixed", "large")
    out = {}
    for name in names:
        if name not in known:
            raise SystemExit(f"--also: unknown workload {name!r}")
        warm = min(args.warmup, 1)
        log(f"other workload: {name} (child process)")
        cmd = [sys.executable, str(Path(__file__).resolve()), "--workload", name, "--cpu-baseline", "0",
               "--warmup", str(warm)]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=900)
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode != 0 or not lines:
            raise SystemExit(f"other workload {name} failed with exit code {r.returncode}")
        d = json.loads(lines[-1])
        rf = d.get("roofline") or {}
        out[name] = {"workload": (d.get("config") or {}).get("workload"), "value": d["value"], "unit": d["unit"],
                     "ms_per_step": d["ms_per_step"], "steps": d["steps"], "warmup": warm,
                     "tflops": rf.get("achieved"), "frac": rf.get("frac"),
                     **({"vs_plain": d["vs_plain"]} if d.get("vs_plain") is not None else {})}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU per process; --gpus must be 1 (sharding across GPUs is not built)", "n_gpus": args.gpus}))
        sys.exit(2)
    log("importing torch")
    import torch

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.workload == "serving":
        out = bench_serving(args)
    elif args.workload == "param_study":
        out = bench_param_study(args)
    elif args.workload == "plain":
        out = bench_plain(args)
    elif args.workload == "env":
        out = bench_env(args)
    elif args.workload == "rollout":
        out = bench_rollout(args)
    else:
        out = bench_ensemble(args, large=args.workload == "large")
    line = {"metric": METRIC, "value": out.pop("value"), "unit": out.pop("unit"), "n_gpus": 1,
            "steps": out.pop("steps"), "warmup": args.warmup, "ms_per_step": out.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic"}
    line.update(out)
    if args.workload == "serving":
        names = [n for n in (OTHER_DEFAULT if args.also is None else args.also).split(",") if n]
        if names:
            line["other_workloads"] = other_workloads(args, names)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
