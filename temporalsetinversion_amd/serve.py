"""Served form: a request loop with continuous batching over a `Batch` (tsi_batch_*).

No counterpart in the reference, whose frame loop (src/main.rs:1056-1087) renders one scene.
A request asks for the next frame of one scene (tape): its t-window is the scene's next 1/60
step, its camera and frame size are the scene's.  `Server.run(requests)` replays an arrival
schedule in real time: requests that have arrived wait in per-scene FIFOs; whenever the device
is free, one submit takes the oldest waiting request of every scene with one (a scene's frames
must run in order: each reuses the octree the previous one left, the temporal reuse of
compute.wgsl:3603-3614), so requests that arrive while a submit runs join the next one.  A
scene's session is opened (admitted by its device footprint against the batch's arena: the HBM
budget) at its first request; a request whose scene cannot be admitted waits until a session
closes (sessions are closed when idle for `idle_close_s`).  Latency = completion - arrival, in
host wall-clock time.

`PlainServer.run(requests)` serves the same schedule with the plain path: one context per scene,
one request at a time in arrival order, each a blocking set inversion and trace (the reference's
one-frame-per-dispatch protocol).  Both leave every scene's octree after the same sequence of
windows, which is how the tests and the benchmark check that a frame is the same alone and in
the mix.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import State, TsiError
from .batch import Batch


@dataclass
class Scene:
    tape: int                  # index into the server's tapes
    localsubdiv: int
    width: int
    height: int
    view: np.ndarray
    params: dict = field(default_factory=dict)   # lipschitz_corrected, ...
    voxel_capacity: int = 1 << 22


@dataclass
class Request:
    scene: int
    arrival: float             # seconds after the start of the run
    frame: int = 0             # the scene's k-th frame: t = k / 60


def poisson_requests(nscenes: int, rate: float, duration: float, seed: int = 0) -> list:
    """Arrivals of a Poisson process of `rate` requests/s over `duration` s, each for a uniformly
    chosen scene; a scene's requests are its consecutive frames."""
    rng = np.random.default_rng(seed)
    t, out, count = 0.0, [], [0] * nscenes
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        s = int(rng.integers(0, nscenes))
        out.append(Request(s, t, count[s]))
        count[s] += 1


def frame_time(k: int) -> float:
    return float(np.float32(k / 60.0))


def _percentiles(lat):
    a = np.asarray(lat, np.float64) * 1e3
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()), "max_ms": float(a.max())}


class Server:
    """Continuous batching over one Batch of the scenes' tapes."""

    def __init__(self, batch: Batch, scenes: list, idle_close_s: float = 1e9):
        self.b = batch
        self.scenes = scenes
        self.sid = {}            # scene -> session
        self.last_used = {}
        self.idle_close_s = idle_close_s
        self.closed = 0          # sessions closed to admit others
        self._waiting = {}
        self._inflight = None

    def _busy(self, s: int) -> bool:
        """The scene has a frame in flight or waiting (its session must stay open)."""
        return bool(self._waiting.get(s)) or any(r.scene == s for r in (self._inflight or []))

    def _admit(self, s: int, now: float) -> bool:
        if s in self.sid:
            return True
        sc = self.scenes[s]
        try:
            self.sid[s] = self.b.open(tape=sc.tape, localsubdiv=sc.localsubdiv, voxel_capacity=sc.voxel_capacity,
                                      max_width=sc.width, max_height=sc.height, **sc.params)
        except TsiError as e:
            if e.code != -10:     # TSI_E_BUDGET: wait for a session to close
                raise
            # close sessions idle for long enough (not while a submit is in flight: closing waits
            # for the device), then retry once
            if self._inflight:
                return False
            for o, t in sorted(self.last_used.items(), key=lambda x: x[1]):
                if o in self.sid and now - t >= self.idle_close_s and not self._busy(o):
                    self.b.close_session(self.sid.pop(o))
                    self.closed += 1
            if s in self.sid:
                return True
            try:
                self.sid[s] = self.b.open(tape=sc.tape, localsubdiv=sc.localsubdiv,
                                          voxel_capacity=sc.voxel_capacity, max_width=sc.width,
                                          max_height=sc.height, **sc.params)
            except TsiError as e2:
                if e2.code != -10:
                    raise
                return False
        return True

    def run(self, requests: list, on_step=None) -> dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        waiting = self._waiting = collections.defaultdict(collections.deque)   # scene -> requests
        backlog = collections.deque()                           # not admitted yet
        done_at = [None] * len(reqs)
        index = {id(r): i for i, r in enumerate(reqs)}
        nxt = 0
        inflight = None
        steps, batch_sizes, set_ms, trace_ms, cells = 0, [], 0.0, 0.0, 0
        failed = 0
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                backlog.append(reqs[nxt])
                nxt += 1
            for _ in range(len(backlog)):
                r = backlog.popleft()
                if self._admit(r.scene, now):
                    waiting[r.scene].append(r)
                else:
                    backlog.append(r)
            if inflight is not None and self.b.done():
                res, s_ms, t_ms = self.b.wait()
                end = time.perf_counter() - t0
                for r, x in zip(inflight, res):
                    done_at[index[id(r)]] = end
                    failed += x["status"] != 0
                    cells += x["cell_evals"]
                set_ms += s_ms
                trace_ms += t_ms
                inflight = self._inflight = None
                if on_step:
                    on_step(res)
            if inflight is None:
                batch = [waiting[s].popleft() for s in list(waiting) if waiting[s]]
                if batch:
                    frames = []
                    for r in batch:
                        sc = self.scenes[r.scene]
                        self.last_used[r.scene] = now
                        frames.append((self.sid[r.scene], frame_time(r.frame), sc.width, sc.height, sc.view))
                    self.b.submit(frames)
                    inflight = self._inflight = batch
                    steps += 1
                    batch_sizes.append(len(batch))
                elif nxt >= len(reqs) and not backlog:
                    break
            if inflight is not None:
                time.sleep(0.0001)
            elif nxt < len(reqs):
                time.sleep(max(0.0, min(0.0005, reqs[nxt].arrival - (time.perf_counter() - t0))))
        lat = [d - r.arrival for d, r in zip(done_at, reqs)]
        span = max(done_at) - reqs[0].arrival if reqs else 0.0
        return dict(_percentiles(lat), requests=len(reqs), steps=steps, frames_per_s=len(reqs) / span if span else 0.0,
                    sessions_closed=self.closed,
                    mean_batch=float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                    max_batch=int(max(batch_sizes)) if batch_sizes else 0, set_inversion_ms=set_ms,
                    trace_ms=trace_ms, cell_evals=cells, failed=failed, span_s=span, latencies=lat)


class PlainServer:
    """The same schedule on the plain path: one context per scene, requests one at a time."""

    def __init__(self, tapes: list, scenes: list):
        self.scenes = scenes
        self.states = {}
        self.tapes = tapes

    def state(self, s: int) -> State:
        if s not in self.states:
            sc = self.scenes[s]
            st = State(self.tapes[sc.tape], width=sc.width, height=sc.height, localsubdiv=sc.localsubdiv,
                       voxel_capacity=sc.voxel_capacity, prune=False, **sc.params)
            st.reset()
            self.states[s] = st
        return self.states[s]

    def run(self, requests: list) -> dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        for s in {r.scene for r in reqs}:      # contexts exist before the clock starts
            self.state(s)
        done_at = []
        failed = 0
        t0 = time.perf_counter()
        for r in reqs:
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            st = self.state(r.scene)
            sc = self.scenes[r.scene]
            st.update(frame_time(r.frame), sc.view)
            try:
                st.set_inversion()
            except TsiError:
                failed += 1
            st.trace(pixels=False)
            st.enqueue_root()
            done_at.append(time.perf_counter() - t0)
        lat = [d - r.arrival for d, r in zip(done_at, reqs)]
        span = max(done_at) - reqs[0].arrival if reqs else 0.0
        return dict(_percentiles(lat), requests=len(reqs), frames_per_s=len(reqs) / span if span else 0.0,
                    failed=failed, span_s=span, latencies=lat)

    def close(self):
        for st in self.states.values():
            st.close()
        self.states = {}
