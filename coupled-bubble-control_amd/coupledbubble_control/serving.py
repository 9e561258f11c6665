"""Serving loop for vector-environment steps (continuous batching).

In the reference an environment step is a blocking call: write the actions
into the control parameters (envs/bubble_1D_pos_control.py:731-749), run
``model.solve`` over *all* envs of the vector env (``step_physics``, :751-753),
then reset time, step size and event value of every env (``advance_time``,
:815-826); terminated envs get new scenes (``reset_envs``, :583-614).

Here many vector envs, of different bubble counts, share one GPU.  Every env
is a system resident in HBM: state and last accepted step size stay in the
device buffers between steps.  A request is one vector env's step: new drive
amplitudes for its envs and K acoustic periods to integrate.

* ``Admission`` (host logic only): requests wait in arrival order; a request
  is admitted when the systems in flight stay within the budget (the memory
  budget divided by the bytes one in-flight system holds).  A request that does
  not fit the budget holds back the ones behind it (no starvation of big
  bursts).  With ``order="venv"`` a request is also admitted only after the
  previous step of the same vector env has finished.
* ``mode="continuous"``: one persistent serving kernel per bubble count
  (``backends.hip.serve.ServeQueue``); admitted requests are appended to its
  work ring and waves with free lanes pick them up while other requests are
  still integrating.  ``order="env"`` (default) orders steps per env, the data
  dependency of a resident system: an env's item of a request is published as
  soon as that env's item of the previous request has finished, so the next
  step of a vector env overlaps the slowest systems of its previous one.
  ``order="venv"`` publishes a request only when the previous step of its
  vector env has finished as a whole (a vector env that blocks on its step).
* ``mode="plain"``: static batching, one launch per batch: when the previous
  batch has finished, every admitted request is written into the device
  buffers and one kernel launch per bubble count integrates them; the batch
  completes when its last launch ends (the reference's blocking ``env.step``
  over everything queued).  ``plain_per_bin=True`` runs one such batch loop per
  bubble count, so the bins' batches overlap each other.
* After an env step completes (``order="env"``) or a request completes
  (``order="venv"`` and the plain path), its envs that did not end with SUCCESS
  (as in ``check_termination``, :771-789) and, with probability ``reset_prob``
  per step, any of its envs start new episodes with fresh radii and positions.
* ``run_async`` (continuous path) drives every env on its own: each env submits
  its next step as soon as its previous one has finished (asynchronous envs, the
  capacity of the serving kernels when no env waits for another).

A system's arithmetic does not depend on which other systems run beside it,
so a request's results are the same alone, in the mix, and in either mode.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np

ITEM_BYTES = 48             # one work-ring item
MISC_BYTES = 32             # tau, h, flag, attempts of a finished item


def inflight_bytes(ups: int, tile: int) -> int:
    """Bytes one in-flight system holds beyond its resident rows: its ring item, its
    host-mapped end-state copy (4 x TILE doubles) and its completion record."""
    return ITEM_BYTES + MISC_BYTES + 4 * tile * 8


@dataclass
class Request:
    rid: int
    n: int                       # bubble count of the vector env
    venv: int                    # vector-env id (its steps are served in order)
    env_ids: np.ndarray          # system indices within the bin of that bubble count
    actions: np.ndarray          # (NK, len(env_ids)) drive amplitudes [bar]
    arrival: float               # seconds after the start of the run
    start: float = -1.0
    done: float = -1.0
    flags: Optional[np.ndarray] = None
    periods: float = 0.0
    seq: Optional[np.ndarray] = None
    attempts: int = -1              # step attempts of its items (continuous mode; -1 unknown)
    left: int = 0                   # envs whose step has not finished (order="env")
    tau: Optional[np.ndarray] = None   # tau reached per env (order="env")
    env_attempts: Optional[np.ndarray] = None   # step attempts per env (order="env")

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def size(self) -> int:
        return int(len(self.env_ids))


class Admission:
    """Arrival-ordered admission under a budget of systems in flight; with
    ``per_venv`` a vector env has at most one request admitted at a time."""

    def __init__(self, max_inflight: int, per_venv: bool = True):
        if max_inflight < 1:
            raise ValueError("max_inflight must be >= 1")
        self.max_inflight = int(max_inflight)
        self.per_venv = bool(per_venv)
        self.pending: Deque[Request] = deque()
        self.inflight = 0
        self.busy_venvs: set = set()

    def push(self, r: Request) -> None:
        if r.size > self.max_inflight:
            raise ValueError(f"request {r.rid} ({r.size} envs) can never fit the budget of {self.max_inflight}")
        self.pending.append(r)

    def admit(self, only_bin: Optional[int] = None) -> List[Request]:
        """Remove and return the requests that may start now (optionally of one bin)."""
        out: List[Request] = []
        keep: Deque[Request] = deque()
        blocked = set(self.busy_venvs)
        room = self.max_inflight - self.inflight
        stop = False
        while self.pending:
            r = self.pending.popleft()
            if stop or (only_bin is not None and r.n != only_bin) or (self.per_venv and r.venv in blocked):
                blocked.add(r.venv)
                keep.append(r)
                continue
            if r.size > room:
                stop = True
                keep.append(r)
                continue
            room -= r.size
            if self.per_venv:
                blocked.add(r.venv)
                self.busy_venvs.add(r.venv)
            out.append(r)
        self.pending = keep
        self.inflight = self.max_inflight - room
        return out

    def finish(self, r: Request) -> None:
        self.inflight -= r.size
        self.busy_venvs.discard(r.venv)

    def release(self, k: int) -> None:
        """k env steps of admitted requests have finished (per-env accounting)."""
        self.inflight -= int(k)


def c0_scene(n: int, k: int, rng, lam: float) -> Tuple[np.ndarray, np.ndarray]:
    """New episodes as in the plain CPU workload: R0 ~ U[20, 80] um, positions sorted
    U[0, 0.25] wavelengths with gaps >= 3 max(R0)."""
    from .workloads import sample_positions

    R0 = rng.uniform(20.0, 80.0, size=(k, n))
    gap = 3.0 * R0.max(axis=1) * 1e-6 / lam
    X0 = sample_positions(rng, k, n, 0.0, 0.25, gap)
    return R0, X0


class ServingEngine:
    """Serves env-step requests over resident CKM1D bins (one model per bubble count)."""

    def __init__(self, bins: Dict[int, object], periods: float = 5.0, mode: str = "continuous",
                 budget_bytes: int = 64 << 20, serve_opts=None, reset_prob: float = 0.0,
                 reset_scene: Optional[Callable] = None, seed: int = 0, keep_step: bool = True,
                 plain_kernel_steps: int = 1 << 20, plain_per_bin: bool = False, order: str = "env"):
        import torch

        from .backends.hip import rkck45_coupled_hip_solver as hs
        from .backends.hip.serve import ServeOptions, ServeQueue

        if mode not in ("continuous", "plain"):
            raise ValueError("mode must be 'continuous' or 'plain'")
        if order not in ("env", "venv"):
            raise ValueError("order must be 'env' or 'venv'")
        # static batches integrate each system once per launch: a vector env's next
        # request always waits for the batch that holds its previous one
        self.order = order if mode == "continuous" else "venv"
        self.torch = torch
        self.hs = hs
        self.bins = dict(bins)
        self.periods = float(periods)
        self.mode = mode
        self.keep_step = keep_step
        self.reset_prob = float(reset_prob)
        self.reset_scene = reset_scene
        self.rng = np.random.default_rng(seed)
        self.plain_kernel_steps = int(plain_kernel_steps)
        self.plain_per_bin = bool(plain_per_bin)
        per_sys = max(inflight_bytes(m._execution_spec.ups, m._execution_spec.tile) for m in self.bins.values())
        self.admission = Admission(max(1, int(budget_bytes) // per_sys), per_venv=self.order == "venv")
        self.queues: Dict[int, object] = {}
        self.streams: Dict[int, object] = {}
        self.pending_reset = {n: np.zeros(m._execution_spec.ns, dtype=bool) for n, m in self.bins.items()}
        for n, m in self.bins.items():
            m._ensure_buffers()
            if m.runtime.solver is None:
                m.runtime.solver = m._make_solver(m.runtime.backend, m.runtime.variant)
            self.streams[n] = torch.cuda.Stream()
            if mode == "continuous":
                o = serve_opts.get(n) if isinstance(serve_opts, dict) else serve_opts
                o = o or ServeOptions()
                # ring slots are freed in submission order, so finished items can wait
                # behind a long-running one: four slots per system that can be in flight,
                # sixteen when envs step on their own (many short steps of other envs
                # then pass one long step)
                cap = 64
                per = 16 if self.order == "env" else 4
                while cap < per * min(self.admission.max_inflight, m._execution_spec.ns):
                    cap *= 2
                o = ServeOptions(**{**o.__dict__, "capacity": min(max(o.capacity, cap), 1 << 24)})
                self.queues[n] = ServeQueue(m, o)
            else:
                # a plain launch integrates every system with actual_time < time_end:
                # systems outside the batch are parked with time_end = actual_time
                db = m.runtime.device_buffers
                db.state.time_end.copy_(db.state.actual_time)
        torch.cuda.synchronize()
        self._init_bookkeeping({n: m._execution_spec.ns for n, m in self.bins.items()})

    def _init_bookkeeping(self, ns_of: Dict[int, int]) -> None:
        """Host-side state of the request loop for bins of ``ns_of[n]`` systems.  Per-env
        bookkeeping of order="env": the request (id) and position of the env's current
        item, its sequence number, whether it is published and unfinished, the env's
        later items in order (and their count), and reserved items not yet published."""
        self.env_req = {n: np.full(ns, -1, dtype=np.int64) for n, ns in ns_of.items()}
        self.env_pos = {n: np.zeros(ns, dtype=np.int64) for n, ns in ns_of.items()}
        self.env_seq = {n: np.zeros(ns, dtype=np.uint32) for n, ns in ns_of.items()}
        self.env_live = {n: np.zeros(ns, dtype=bool) for n, ns in ns_of.items()}
        self.env_wait: Dict[int, Dict[int, Deque[Tuple[Request, int]]]] = {n: {} for n in ns_of}
        self.wait_count = {n: np.zeros(ns, dtype=np.int64) for n, ns in ns_of.items()}
        self.ready: Dict[int, List[Tuple[Request, np.ndarray]]] = {n: [] for n in ns_of}
        self.live: Dict[int, Request] = {}
        self.inflight: List[Request] = []
        self.batches: Dict[int, Tuple[object, List[Request]]] = {}
        self.completed: List[Request] = []
        self.resets = 0
        self.ring_full = 0
        self._t0 = 0.0
        self._feeder = None

    # -- helpers ---------------------------------------------------------------
    def _now(self) -> float:
        return time.perf_counter() - self._t0

    def _h0(self) -> float:
        return 0.0 if self.keep_step else 1e-6

    def _finish(self, r: Request, flags: np.ndarray, t_reached: np.ndarray, now: float) -> None:
        r.flags = flags.astype(np.int32)
        r.periods = float(t_reached.sum())
        r.done = now
        self.admission.finish(r)
        self.completed.append(r)
        self._reset_after(r)
        if self._feeder is not None:
            for nr in self._feeder(r, now):
                self.admission.push(nr)

    def _reset_after(self, r: Request) -> None:
        if self.reset_scene is None:
            return
        mask = r.flags != 1
        if self.reset_prob > 0.0:
            mask |= self.rng.random(r.size) < self.reset_prob
        if not mask.any():
            return
        ids = np.asarray(r.env_ids)[mask]
        m = self.bins[r.n]
        R0, X0 = self.reset_scene(r.n, ids.size, self.rng)
        if self.mode == "continuous":
            # no device work beside the serving kernels: the new episode's rows are
            # staged in host memory and taken by the env's next item
            state, up, cm = m.scene_rows(ids, R0, X0)
            self.queues[r.n].stage(ids, state, up, cm)
            self.pending_reset[r.n][ids] = True
        else:
            st = self.streams[r.n]
            m.reset_systems(ids, R0, X0, h0=1e-6, stream=st)
            with self.torch.cuda.stream(st):
                ti = self.torch.from_numpy(ids.astype(np.int64)).to(m.runtime.device_buffers.state.time_end.device)
                m.runtime.device_buffers.state.time_end.index_fill_(0, ti, 0.0)
        self.resets += int(ids.size)

    # -- continuous ------------------------------------------------------------
    def _submit(self, reqs: List[Request], now: float) -> None:
        by_bin: Dict[int, List[Request]] = {}
        for r in reqs:
            by_bin.setdefault(r.n, []).append(r)
        for n, rs in by_bin.items():
            q = self.queues[n]
            ids = np.concatenate([r.env_ids for r in rs]).astype(np.int32)
            pa = np.concatenate([r.actions for r in rs], axis=1) * 1.0e5
            reset = self.pending_reset[n][ids]
            h0 = np.where(reset, 1e-6, self._h0())
            seq = q.submit(ids, pa, self.periods, h0, reset if reset.any() else None)
            if seq is None:
                # ring slots still held by finished items behind a long one: retry later,
                # keeping arrival order
                for r in reversed(rs):
                    self.admission.finish(r)
                    self.admission.pending.appendleft(r)
                self.ring_full += 1
                continue
            self.pending_reset[n][ids] = False
            off = 0
            for r in rs:
                r.seq = seq[off:off + r.size]
                r.start = now
                off += r.size
            self.inflight.extend(rs)

    def _collect(self, now: float) -> int:
        keep, n_done = [], 0
        for r in self.inflight:
            q = self.queues[r.n]
            if q.done(r.env_ids, r.seq).all():
                _, misc = q.results(r.env_ids)
                r.attempts = int(misc[:, 3].sum())
                self._finish(r, misc[:, 2], misc[:, 0], now)
                n_done += 1
            else:
                keep.append(r)
        self.inflight = keep
        return n_done

    # -- continuous, per-env order ----------------------------------------------------
    def _reset_envs(self, n: int, ids: np.ndarray, flags: np.ndarray) -> None:
        """New episodes for finished envs ``ids`` of bin n that failed (flag != SUCCESS)
        or are drawn with ``reset_prob``: their rows are staged for their next item."""
        if self.reset_scene is None or ids.size == 0:
            return
        mask = flags != 1
        if self.reset_prob > 0.0:
            mask |= self.rng.random(ids.size) < self.reset_prob
        if not mask.any():
            return
        ids = ids[mask]
        R0, X0 = self.reset_scene(n, ids.size, self.rng)
        state, up, cm = self.bins[n].scene_rows(ids, R0, X0)
        self.queues[n].stage(ids, state, up, cm)
        self.pending_reset[n][ids] = True
        self.resets += int(ids.size)

    def _take(self, n: int, r: Request, pos: np.ndarray) -> None:
        """Reserve the envs at positions ``pos`` of request r for their next item."""
        ids = r.env_ids[pos]
        self.env_req[n][ids] = r.rid
        self.env_pos[n][ids] = pos
        self.ready[n].append((r, pos))

    def _admit_env(self, reqs: List[Request], now: float) -> None:
        """Admitted requests, in arrival order: envs that are idle (no item running and
        none queued before this one) are reserved for publication, the others queue
        behind their env's earlier steps."""
        for r in reqs:
            n = r.n
            r.start = now
            r.left = r.size
            r.flags = np.zeros(r.size, dtype=np.int32)
            r.tau = np.zeros(r.size)
            r.env_attempts = np.zeros(r.size, dtype=np.int64)
            r.attempts = 0
            self.live[r.rid] = r
            ids = np.asarray(r.env_ids)
            wait = self.env_wait[n]
            free = (self.env_req[n][ids] < 0) & (self.wait_count[n][ids] == 0)
            busy = np.flatnonzero(~free)
            for p in busy.tolist():
                wait.setdefault(int(ids[p]), deque()).append((r, p))
            self.wait_count[n][ids[busy]] += 1
            pos = np.flatnonzero(free)
            if pos.size:
                self._take(n, r, pos)

    def _publish(self, now: float) -> None:
        """Publish every bin's reserved items in one submission (kept for the next round
        when the ring has too few free slots)."""
        for n, items in self.ready.items():
            if not items:
                continue
            ids = np.concatenate([np.asarray(r.env_ids)[p] for r, p in items]).astype(np.int32)
            pa = np.concatenate([r.actions[:, p] for r, p in items], axis=1) * 1.0e5
            reset = self.pending_reset[n][ids]
            h0 = np.where(reset, 1e-6, self._h0())
            seq = self.queues[n].submit(ids, pa, self.periods, h0, reset if reset.any() else None)
            if seq is None:
                self.ring_full += 1
                continue
            self.pending_reset[n][ids] = False
            self.env_seq[n][ids] = seq
            self.env_live[n][ids] = True
            self.ready[n] = []

    def _collect_env(self, now: float) -> int:
        """Finished env items: record their results in their requests, start new episodes,
        reserve each env's next queued item, and complete requests whose envs are all done."""
        n_done = 0
        for n, q in self.queues.items():
            live = np.flatnonzero(self.env_live[n])
            if live.size == 0:
                continue
            fin = live[q.done(live, self.env_seq[n][live])]
            if fin.size == 0:
                continue
            misc = q.out_misc[fin].copy()          # [tau, h, flag, attempts]
            self.env_live[n][fin] = False
            rids = self.env_req[n][fin]
            pos = self.env_pos[n][fin]
            self.env_req[n][fin] = -1
            self.admission.release(fin.size)
            finished = []
            for rid in np.unique(rids):
                sel = rids == rid
                r = self.live[int(rid)]
                r.flags[pos[sel]] = misc[sel, 2].astype(np.int32)
                r.tau[pos[sel]] = misc[sel, 0]
                r.env_attempts[pos[sel]] = misc[sel, 3].astype(np.int64)
                r.attempts += int(misc[sel, 3].sum())
                r.left -= int(sel.sum())
                if r.left == 0:
                    finished.append(r)
            self._reset_envs(n, fin, misc[:, 2])
            wq = fin[self.wait_count[n][fin] > 0]
            if wq.size:
                wait = self.env_wait[n]
                nxt: Dict[int, Tuple[Request, List[int]]] = {}
                for e in wq.tolist():
                    dq = wait[e]
                    r2, p2 = dq.popleft()
                    if not dq:
                        del wait[e]
                    nxt.setdefault(r2.rid, (r2, []))[1].append(p2)
                self.wait_count[n][wq] -= 1
                for r2, ps in nxt.values():
                    self._take(n, r2, np.array(ps, dtype=np.int64))
            for r in sorted(finished, key=lambda x: x.rid):
                del self.live[r.rid]
                r.periods = float(r.tau.sum())
                r.done = now
                self.completed.append(r)
                n_done += 1
                if self._feeder is not None:
                    for nr in self._feeder(r, now):
                        self.admission.push(nr)
        return n_done

    def run_async(self, duration: float, seed: int = 0, pa_max: float = 1.5, timeout_s: float = 600.0,
                  progress: Optional[Callable[[str], None]] = None) -> Dict[str, float]:
        """Asynchronous envs (continuous path): every resident env submits a step with new
        amplitudes ~ U[0, pa_max] bar at t = 0 and, until ``duration`` seconds have passed,
        its next step as soon as its previous one has finished; episode resets as in
        ``run``.  Returns throughput and per-env-step latency (submission to completion)."""
        if self.mode != "continuous":
            raise ValueError("run_async drives the continuous path")
        rng = np.random.default_rng(seed)
        nk = {n: m._model_spec.nk for n, m in self.bins.items()}
        start = {n: np.zeros(m._execution_spec.ns) for n, m in self.bins.items()}
        lat: List[np.ndarray] = []
        periods, steps, failed = 0.0, 0, 0
        self._t0 = time.perf_counter()
        last_report = 0.0

        def submit(n: int, ids: np.ndarray, now: float) -> None:
            pa = rng.uniform(0.0, pa_max, size=(nk[n], ids.size)) * 1.0e5
            reset = self.pending_reset[n][ids]
            h0 = np.where(reset, 1e-6, self._h0())
            while True:
                seq = self.queues[n].submit(ids.astype(np.int32), pa, self.periods, h0,
                                            reset if reset.any() else None)
                if seq is not None:
                    break
                self.ring_full += 1
                self.queues[n].poll()
                time.sleep(2e-5)
            self.pending_reset[n][ids] = False
            self.env_seq[n][ids] = seq
            self.env_live[n][ids] = True
            start[n][ids] = now

        for n, m in self.bins.items():
            submit(n, np.arange(m._execution_spec.ns), 0.0)
        while any(v.any() for v in self.env_live.values()):
            now = self._now()
            if now > timeout_s:
                raise TimeoutError(f"asynchronous run exceeded {timeout_s} s")
            moved = 0
            for n, q in self.queues.items():
                live = np.flatnonzero(self.env_live[n])
                if live.size == 0:
                    continue
                fin = live[q.done(live, self.env_seq[n][live])]
                if fin.size == 0:
                    continue
                misc = q.out_misc[fin].copy()
                self.env_live[n][fin] = False
                lat.append(now - start[n][fin])
                periods += float(misc[:, 0].sum())
                steps += int(fin.size)
                failed += int((misc[:, 2] != 1).sum())
                self._reset_envs(n, fin, misc[:, 2])
                if now < duration:
                    submit(n, fin, now)
                moved += fin.size
            for q in self.queues.values():
                q.poll()
            if not moved:
                time.sleep(2e-5)
            if progress and now - last_report > 5.0:
                last_report = now
                progress(f"serving[async] t={now:.1f}s env steps {steps}")
        wall = self._now()
        ms = np.concatenate(lat) * 1e3 if lat else np.zeros(1)
        return {"mode": "continuous_async", "wall_s": wall, "env_steps": steps, "periods": periods,
                "periods_per_s": periods / wall, "env_steps_per_s": steps / wall,
                "p50_ms": float(np.percentile(ms, 50)), "p99_ms": float(np.percentile(ms, 99)),
                "mean_ms": float(ms.mean()), "max_ms": float(ms.max()), "failed_env_steps": failed,
                "resets": self.resets, "ring_full_retries": self.ring_full}

    # -- plain -------------------------------------------------------------------
    def _launch_batch(self, n: int, reqs: List[Request], now: float) -> None:
        torch = self.torch
        m = self.bins[n]
        db = m.runtime.device_buffers
        st = self.streams[n]
        nk = m._model_spec.nk
        ids = np.concatenate([r.env_ids for r in reqs]).astype(np.int64)
        pa = np.concatenate([r.actions for r in reqs], axis=1) * 1.0e5
        dev = db.state.actual_state.device
        with torch.cuda.stream(st):
            ti = torch.from_numpy(ids).to(dev)
            a = torch.from_numpy(np.ascontiguousarray(pa)).to(dev)
            cp = db.params.control_params
            for k in range(nk):
                cp[k].index_copy_(0, ti, a[k])
            db.state.actual_time.index_fill_(0, ti, 0.0)
            db.state.time_end.index_fill_(0, ti, self.periods)
            db.status.actual_event.index_fill_(0, ti, -1.0)
            db.status.status_flags.index_fill_(0, ti, 0)
            if not self.keep_step:
                db.state.time_step.index_fill_(0, ti, 1e-6)
            self.hs.launch(m.runtime.kernel, db, self.plain_kernel_steps, stream=st)
            ev = torch.cuda.Event()
            ev.record(st)
        for r in reqs:
            r.start = now
        self.batches[n] = (ev, reqs, ti)

    def _collect_plain(self, now: float) -> int:
        """Complete finished launches: per bin, or the whole batch once all its
        launches have ended."""
        ready = [n for n, (ev, _, _) in self.batches.items() if ev.query()]
        if not self.plain_per_bin and len(ready) < len(self.batches):
            return 0
        n_done = 0
        for n in ready:
            ev, reqs, ti = self.batches.pop(n)
            db = self.bins[n].runtime.device_buffers
            with self.torch.cuda.stream(self.streams[n]):
                flags = db.status.status_flags.index_select(0, ti).cpu().numpy()
                t_now = db.state.actual_time.index_select(0, ti)
                t = t_now.cpu().numpy()
                db.state.time_end.index_copy_(0, ti, t_now)   # park the batch again
            off = 0
            for r in reqs:
                self._finish(r, flags[off:off + r.size], t[off:off + r.size], now)
                off += r.size
                n_done += 1
        return n_done

    # -- main loop -------------------------------------------------------------------
    def run(self, requests: Sequence[Request], timeout_s: float = 600.0,
            progress: Optional[Callable[[str], None]] = None,
            feeder: Optional[Callable[[Request, float], List[Request]]] = None) -> Dict[str, float]:
        """Serve the requests (arrival times relative to the start of the run) until all
        have completed; ``feeder(done_request, now)`` may return follow-up requests, which
        arrive at once (a closed loop).  Returns throughput and latency statistics."""
        arrivals = deque(sorted(requests, key=lambda r: r.arrival))
        self.completed = []
        self._feeder = feeder
        self._t0 = time.perf_counter()
        last_report = 0.0
        while arrivals or self.admission.pending or self.inflight or self.batches or self.live:
            now = self._now()
            if now > timeout_s:
                raise TimeoutError(f"serving run exceeded {timeout_s} s ({len(self.completed)} requests done)")
            moved = 0
            while arrivals and arrivals[0].arrival <= now:
                self.admission.push(arrivals.popleft())
                moved += 1
            if self.mode == "continuous" and self.order == "env":
                done = self._collect_env(now)
                adm = self.admission.admit()
                if adm:
                    self._admit_env(adm, now)
                self._publish(now)
                for q in self.queues.values():
                    q.poll()
            elif self.mode == "continuous":
                done = self._collect(now)
                adm = self.admission.admit()
                if adm:
                    self._submit(adm, now)
                for q in self.queues.values():
                    q.poll()
            else:
                done = self._collect_plain(now)
                adm = []
                if self.plain_per_bin:
                    for n in self.bins:
                        if n not in self.batches:
                            a = self.admission.admit(only_bin=n)
                            if a:
                                self._launch_batch(n, a, now)
                                adm += a
                elif not self.batches:
                    adm = self.admission.admit()
                    by_bin: Dict[int, List[Request]] = {}
                    for r in adm:
                        by_bin.setdefault(r.n, []).append(r)
                    for n, rs in by_bin.items():
                        self._launch_batch(n, rs, now)
            if not (moved or done or adm):
                time.sleep(2e-5)
            if progress and now - last_report > 5.0:
                last_report = now
                progress(f"serving[{self.mode}] t={now:.1f}s done {len(self.completed)}")
        wall = self._now()
        self._feeder = None
        return self.summary(wall)

    def summary(self, wall: float) -> Dict[str, float]:
        lat = np.array([r.latency for r in self.completed]) * 1e3
        env_steps = sum(r.size for r in self.completed)
        periods = sum(r.periods for r in self.completed)
        failed = sum(int((r.flags != 1).sum()) for r in self.completed)
        mode = self.mode + ("_per_bin" if self.mode == "plain" and self.plain_per_bin else "")
        if self.mode == "continuous":
            mode += "_" + self.order
        out = {"mode": mode, "wall_s": wall, "requests": len(self.completed), "env_steps": env_steps,
               "periods": periods, "periods_per_s": periods / wall, "env_steps_per_s": env_steps / wall,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean()), "max_ms": float(lat.max()), "failed_env_steps": failed,
               "resets": self.resets}
        if self.mode == "continuous":
            out["ring_full_retries"] = self.ring_full
            out["kernel_launches"] = int(sum(q.stats()[0] for q in self.queues.values()))
        return out

    def window(self, lo: int, hi: int) -> Dict[str, float]:
        """Statistics of the requests with id in [lo, hi) (a steady-state window: requests
        before lo load the queues, requests after hi keep them loaded while the window's
        requests finish).  The window runs from the first of its requests' arrivals to the
        last of their completions; throughput is their integrated periods over that span,
        latency is per request (arrival to completion)."""
        rs = [r for r in self.completed if lo <= r.rid < hi]
        if len(rs) != hi - lo:
            raise ValueError(f"window [{lo}, {hi}) has {len(rs)} completed requests")
        t0 = min(r.arrival for r in rs)
        t1 = max(r.done for r in rs)
        wall = t1 - t0
        lat = np.array([r.latency for r in rs]) * 1e3
        periods = sum(r.periods for r in rs)
        env_steps = sum(r.size for r in rs)
        att: Dict[int, int] = {}
        for r in rs:
            att[r.n] = att.get(r.n, 0) + max(r.attempts, 0)
        return {"wall_s": wall, "requests": len(rs), "env_steps": env_steps, "periods": periods,
                "periods_per_s": periods / wall, "env_steps_per_s": env_steps / wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
                "failed_env_steps": sum(int((r.flags != 1).sum()) for r in rs),
                "attempts": att if all(r.attempts >= 0 for r in rs) else None}

    def close(self) -> None:
        for q in self.queues.values():
            q.close()
        self.queues = {}


def vector_envs(bin_sizes: Dict[int, int], rng, lo: int = 64, hi: int = 1024) -> List[Tuple[int, np.ndarray]]:
    """Partition each bin's systems into vector envs of integers(lo, hi + 1) envs (the last
    one of a bin takes the remainder, merged into its predecessor if below lo)."""
    out = []
    for n, ns in bin_sizes.items():
        sizes = []
        left = ns
        while left > 0:
            k = int(rng.integers(lo, hi + 1))
            k = min(k, left)
            sizes.append(k)
            left -= k
        if len(sizes) > 1 and sizes[-1] < lo:
            last = sizes.pop()
            sizes[-1] += last
        off = 0
        for k in sizes:
            out.append((n, np.arange(off, off + k, dtype=np.int32)))
            off += k
    return out


def make_requests(venvs: Sequence[Tuple[int, np.ndarray]], rate_env_steps: float, duration: Optional[float] = None,
                  seed: int = 0, pa_max: float = 1.5, nk: int = 2, count: Optional[int] = None) -> List[Request]:
    """Poisson arrivals of vector-env steps at ``rate_env_steps`` env-steps/s in aggregate
    (bursts at rate / mean vector-env size per second) for ``duration`` seconds or
    ``count`` requests; the vector envs take turns in a shuffled round-robin order;
    actions ~ U[0, pa_max] bar."""
    if (duration is None) == (count is None):
        raise ValueError("give exactly one of duration and count")
    rng = np.random.default_rng(seed)
    mean_size = float(np.mean([len(ids) for _, ids in venvs]))
    lam = rate_env_steps / mean_size
    order = rng.permutation(len(venvs))
    t, j, out = 0.0, 0, []
    while True:
        t += rng.exponential(1.0 / lam)
        if (duration is not None and t > duration) or (count is not None and len(out) >= count):
            break
        v = int(order[j % len(order)])
        j += 1
        n, ids = venvs[v]
        out.append(Request(len(out), n, v, ids, rng.uniform(0.0, pa_max, size=(nk, len(ids))), t))
    return out


def closed_loop(venvs: Sequence[Tuple[int, np.ndarray]], duration: float, seed: int = 0, pa_max: float = 1.5,
                nk: int = 2):
    """Requests for a closed loop: every vector env submits at t = 0 and, until ``duration``
    seconds have passed, its next step as soon as the previous one completes (the
    blocking env.step of a PPO rollout worker).  Returns (initial requests, feeder)."""
    rng = np.random.default_rng(seed)
    counter = [0]

    def new(v: int, t: float) -> Request:
        n, ids = venvs[v]
        r = Request(counter[0], n, v, ids, rng.uniform(0.0, pa_max, size=(nk, len(ids))), t)
        counter[0] += 1
        return r

    first = [new(v, 0.0) for v in range(len(venvs))]

    def feeder(done: Request, now: float) -> List[Request]:
        return [new(done.venv, now)] if now < duration else []

    return first, feeder
