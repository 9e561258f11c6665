"""Ensembles of systems with different bubble counts (the batched uneven workload).

The reference compiles one kernel per bubble count (UPS is a compile-time macro,
backends/cupy_cuda/kernel_loader.py:49) and can integrate only one bubble count
per model.  Here an ensemble holds one CKM1D model per bubble count ("bin", each
in the reference buffer layout) and integrates all of them together:

* each bin's systems go through a persistent work queue: lanes or lane groups
  take that bin's systems longest-horizon-first and refill as systems stop, so
  uneven horizons and step counts do not idle lanes;
* when every bin uses lane groups and a fused kernel covers the bubble counts
  ({8, 4, 3, 2} or {16, 2}), all bins run in ONE persistent kernel
  (cbc_launch_fused): each wave drains the bins' queues in turn, largest bubble
  count first, so the longest attempt chains start first and waves that find a
  bin drained move on to the next.  The step of the 32 768-system mixed set is
  its slowest N = 8 system's chain either way (5.04 s fused, 5.00 s with the
  per-bin launches below, under the profiler); the bins' results are bit for
  bit those of their own launches;
* otherwise (or with fused="off") each bin runs one persistent launch
  (cbc_launch_queued), issued on side streams without host
  synchronisation between them, so when one bin's queue drains its waves free
  SIMDs for the others (the tails overlap).  A bin's time is bounded below by
  its slowest system's chain of attempts, which grows with N, so the largest
  bubble count is issued first, alone on the first side stream, and the other
  bins go round-robin, largest first, onto at most two more (queued launches
  on one stream run one after the other).  At most three side streams: with
  the current stream they fit the four hardware queues of the device, and a
  launch whose stream shares a hardware queue with another bin's waits for
  that launch to end.  32 768-system mixed set, in an earlier build with slower
  N = 8 attempts: 8.0-8.2 s this way, 11.7-12.0 s with one stream per bin and the
  N = 8 bin on the fourth, the N = 8 bin alone 7.8 s;
* the launch sequence can be captured once into a HIP graph and replayed;
* ``autotune`` picks, per bin, the mapping (lane vs group), the refill
  threshold and the lane mapping's workgroup size from short timed trials.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import torch

from .backends import KernelVariant
from .backends.hip import rkck45_coupled_hip_solver as hs


MAX_SIDE_STREAMS = 3   # hardware queues per device (4) minus the current stream's


def stream_slots(bubble_counts: Sequence[int], streams: int) -> List[tuple]:
    """Issue order and side stream of each bin: (bin index, stream index) pairs,
    largest bubble count first; the first bin alone on stream 0, the others
    round-robin over the remaining streams (all on stream 0 when there is one)."""
    if streams < 1:
        raise ValueError("need at least one stream")
    order = sorted(range(len(bubble_counts)), key=lambda i: -bubble_counts[i])
    return [(i, 0 if j == 0 or streams == 1 else 1 + (j - 1) % (streams - 1)) for j, i in enumerate(order)]


@dataclass
class BinTuning:
    """Launch choice for one bin."""
    variant: str = "auto"      # "lane" | "group" | "auto"
    refill: int = 1            # free lanes (groups) per wave that trigger a refill
    waves: int = 0             # persistent grid in waves (0 = fill every SIMD)
    block: int = 0             # threads per workgroup of the lane mapping (0 = 64)


@dataclass
class TuneRecord:
    n: int
    choice: BinTuning
    trials: Dict[str, float] = field(default_factory=dict)


FUSED_MODES = ("auto", "on", "off")


class MixedEnsemble:
    """One CKM1D model per bubble count, integrated together.  ``fused``: "on" runs
    every bin in one fused persistent kernel (all bins in the group mapping, whatever
    their tuning says; an error if no fused kernel covers the bubble counts), "off"
    runs one queued launch per bin on side streams, "auto" fuses two or more bins when
    a fused kernel covers them and every bin's plan uses the group mapping."""

    def __init__(self, models: Sequence, tuning: Optional[Dict[int, BinTuning]] = None, fused: str = "auto"):
        self.models = list(models)
        ns = [m._execution_spec.ups for m in self.models]
        if len(set(ns)) != len(ns):
            raise ValueError("one model per bubble count")
        if fused not in FUSED_MODES:
            raise ValueError(f"fused must be one of {FUSED_MODES}, got {fused!r}")
        self.tuning = {m._execution_spec.ups: (tuning or {}).get(m._execution_spec.ups, BinTuning())
                       for m in self.models}
        self.fused = fused
        self.fused_active = False
        self._streams: List[torch.cuda.Stream] = []
        self._orders: List[torch.Tensor] = []
        self._graph = None

    # -- set-up ---------------------------------------------------------------
    def _fusable(self) -> bool:
        ms = self.models[0]._model_spec
        return (len({(m._model_spec.nk, m._model_spec.ac) for m in self.models}) == 1
                and hs.fused_supported([m._execution_spec.ups for m in self.models], ms.nk, ms.ac))

    def _prepare(self):
        # one bin alone runs its own kernel (built with that bubble count's flags)
        fuse = (self.fused == "on" or (self.fused == "auto" and len(self.models) > 1)) and self._fusable()
        if self.fused == "on" and not fuse:
            raise ValueError("no fused kernel covers bubble counts "
                             f"{sorted(m._execution_spec.ups for m in self.models)} at these specs")
        for m in self.models:
            m._ensure_buffers()
            t = self.tuning[m._execution_spec.ups]
            want = KernelVariant("group" if self.fused == "on" else t.variant)
            block = None if want == KernelVariant.GROUP else (t.block or None)
            if m.runtime.kernel is None or m.runtime.variant != want or m._hip_opts.block_size != block:
                m.runtime.variant = want
                m._hip_opts.block_size = block
                m.runtime.solver = m._make_solver(m.runtime.backend, want)
        self.fused_active = fuse and all(m.runtime.kernel.info().mapping == "group" for m in self.models)
        k = 0 if self.fused_active else min(len(self.models), MAX_SIDE_STREAMS)
        if len(self._streams) != k:
            self._streams = [torch.cuda.Stream() for _ in range(k)]

    def orders(self) -> List[torch.Tensor]:
        """Longest-remaining-horizon-first queues of every bin (kept alive for replays)."""
        self._orders = [hs.horizon_order(m.runtime.device_buffers) for m in self.models]
        return self._orders

    # -- execution ------------------------------------------------------------
    def launch(self, kernel_steps: int = 1 << 20, orders: Optional[List[torch.Tensor]] = None) -> None:
        """Issue the bins' persistent launches (asynchronous; capturable): one fused
        kernel on the current stream, or one launch per bin on side streams joined back
        into the current stream."""
        self._prepare()
        orders = orders if orders is not None else self.orders()
        main = torch.cuda.current_stream()
        if self.fused_active:
            live = [i for i, o in enumerate(orders) if o.numel() > 0]
            if live:
                ms = [self.models[i] for i in live]
                hs.launch_fused([m.runtime.kernel for m in ms], [m.runtime.device_buffers for m in ms],
                                [orders[i] for i in live],
                                [self.tuning[m._execution_spec.ups].refill for m in ms],
                                waves=0, kernel_steps=kernel_steps, stream=main)
            return
        for st in self._streams:
            st.wait_stream(main)
        ups = [m._execution_spec.ups for m in self.models]
        for i, slot in stream_slots(ups, len(self._streams)):
            m, order, st = self.models[i], orders[i], self._streams[slot]
            if order.numel() == 0:
                continue
            t = self.tuning[m._execution_spec.ups]
            with torch.cuda.stream(st):
                hs.launch_queued(m.runtime.kernel, m.runtime.device_buffers, order, waves=t.waves,
                                 refill=t.refill, kernel_steps=kernel_steps, stream=st)
        for st in self._streams:
            main.wait_stream(st)

    def solve(self, kernel_steps: int = 1 << 20, max_launches: int = 64) -> int:
        """Launch until no system of any bin is ACTIVE; returns the launch rounds.  Rounds
        after the first queue only the systems still ACTIVE, so a system that stopped on a
        failure (STEP_SIZE_LIMIT, NaN) stays stopped, as in the reference host loop."""
        rounds = 0
        orders = None
        for _ in range(max_launches):
            self.launch(kernel_steps, orders=orders)
            rounds += 1
            main = torch.cuda.current_stream()
            main.synchronize()   # joined every side stream in launch()
            left = sum(hs.active_count(m.runtime.kernel, main) for m in self.models)
            if left == 0:
                break
            orders = [hs.active_horizon_order(m.runtime.device_buffers) for m in self.models]
            self._orders = orders
        return rounds

    def capture(self, kernel_steps: int = 1 << 20):
        """Capture one launch round into a HIP graph (replayed by ``replay``).  The
        queue order is fixed at capture time."""
        self._prepare()
        orders = self.orders()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch(kernel_steps, orders=orders)
        self._graph = g
        return g

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()

    # -- results --------------------------------------------------------------
    def integrated_periods(self) -> float:
        return float(sum((m.runtime.device_buffers.state.actual_time.sum()).item() for m in self.models))

    def attempts(self) -> Dict[int, int]:
        return {m._execution_spec.ups: int(m.runtime.device_buffers.status.total_steps.to(torch.int64).sum().item())
                for m in self.models}


def _trial_copy(m, horizon: float, limit: int):
    """A bin copy of at most `limit` systems integrating `horizon` periods."""
    import numpy as np

    from .models.ckm1d import CKM1D

    es = m._execution_spec
    ns = min(es.ns, limit)
    hb = m.runtime.host_buffers
    R0 = m._R0[:ns, :es.ups] * 1e6
    c = CKM1D(num_systems=ns, systems_per_block=64, units_per_system=es.ups, R0=R0,
              PA=m._PA[:, :ns] / 1e5, FR=m._FREQ[:, :ns] / 1e3, PS=m._PS[:, :ns] / np.pi,
              REF_FREQ=m._REF_FREQ / 1e3, num_components=m._model_spec.nk, acoustic_field=m._model_spec.ac,
              mat_props=m.mat_props, solver_spec=m._solver_spec)
    c.runtime.host_buffers.state.actual_state[:] = hb.state.actual_state[:, :ns]
    c.runtime.host_buffers.state.time_end[:] = np.minimum(hb.state.time_end[:ns], horizon)
    return c


def autotune(models: Sequence, horizon: float = 0.5, limit: int = 8192,
             refills: Sequence[int] = (1, 16, 64), blocks: Sequence[int] = (64, 256),
             cache: Optional[Path] = None) -> Dict[int, TuneRecord]:
    """Pick mapping, refill threshold and (lane mapping) workgroup size per bin from timed
    trials of `horizon` periods on up to `limit` systems of that bin.  The lane mapping is
    tried for N <= 2 only (it spills beyond).  Results are cached as JSON keyed by bubble
    count when `cache` is given."""
    if cache is not None and Path(cache).exists():
        data = json.loads(Path(cache).read_text())
        return {int(k): TuneRecord(v["n"], BinTuning(**v["choice"]), v["trials"]) for k, v in data.items()}
    out: Dict[int, TuneRecord] = {}
    for m in models:
        n = m._execution_spec.ups
        cands = [BinTuning("group", rf, 0, 0) for rf in refills]
        if n <= 2:
            cands = [BinTuning("lane", rf, 0, b) for b in blocks for rf in refills] + cands
        best, best_t, trials = None, float("inf"), {}
        for cand in cands:
            c = _trial_copy(m, horizon, limit)
            ens = MixedEnsemble([c], {n: cand})
            ens._prepare()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ens.solve()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            key = f"{cand.variant}/refill={cand.refill}" + (f"/block={cand.block}" if cand.block else "")
            trials[key] = dt
            if dt < best_t:
                best, best_t = cand, dt
        out[n] = TuneRecord(n, best, trials)
    if cache is not None:
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps({str(k): {"n": v.n, "choice": asdict(v.choice), "trials": v.trials}
                                           for k, v in out.items()}, indent=1))
    return out
