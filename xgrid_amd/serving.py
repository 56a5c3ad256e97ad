"""Continuous-batching serving loop for stencil solves.

No counterpart in the reference.  Requests (``workloads.InstanceSpec`` with an
arrival time and a step count) arrive over time; the server

* admits queued requests in arrival order while their device footprint fits
  the HBM budget and their cells fit the resident-cell cap (a request waits at
  the head of the queue until enough memory and cells are free).  The cap is
  how much work one batched launch needs to keep the GPU busy: admitting more
  adds no throughput, it only makes every resident request take longer;
* keeps every admitted request's fields on the device and advances all of
  them together, a few kernel calls per tick, in one batch per kernel
  (``multi.MultiSolver``): requests that arrive while others run join the
  batch at the next tick, so their work overlaps on the device;
* retires a request when its steps are done, copies its fields to the host
  (the response) and frees its memory.

Latency = response ready - arrival.  ``serve_plain`` runs the same trace the
drop-in way (one request at a time, plain kernel calls) for comparison.
Each request's result equals the same request solved alone (the batched
device code per cell is the plain code).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .multi import MultiSolver
from .workloads import InstanceSpec


def footprint(spec: InstanceSpec, elem_bytes: int) -> int:
    """Device bytes of one request: per grid two time levels and one scratch of
    fields, plus the int32 mask (a cavity has four grids, diffusion one)."""
    grids = 4 if spec.kind == "cavity" else 1
    return grids * spec.n * spec.n * (3 * elem_bytes + 4)


class _DeviceTimer:
    """Device time spent in serving ticks: an event before and after each tick's
    work on the backend stream (the gaps between ticks are not counted)."""

    def __init__(self):
        self.lib = _lib.load()
        self.pending = []
        self.total_ms = 0.0

    def begin(self):
        self._a = self._record()

    def end(self):
        self.pending.append((self._a, self._record()))
        if len(self.pending) > 64:
            self._drain(len(self.pending) - 32)

    def sync(self):
        _lib.check(self.lib.xg_synchronize(), "xg_synchronize")

    def mark(self):
        """An event after the work queued so far (release it with `release`)."""
        return self._record()

    def wait(self, event):
        """Wait until the work before `event` is done."""
        import ctypes
        ms = ctypes.c_double()
        _lib.check(self.lib.xg_event_elapsed(event, event, ctypes.byref(ms)), "xg_event_elapsed")

    def release(self, event):
        self.lib.xg_event_destroy(event)

    def busy_seconds(self) -> float:
        self._drain(len(self.pending))
        t, self.total_ms = self.total_ms, 0.0
        return t / 1e3

    def _record(self):
        import ctypes
        e = ctypes.c_void_p()
        _lib.check(self.lib.xg_event_record(ctypes.byref(e)), "xg_event_record")
        return e.value

    def _drain(self, n: int):
        import ctypes
        for a, b in self.pending[:n]:
            ms = ctypes.c_double()
            _lib.check(self.lib.xg_event_elapsed(a, b, ctypes.byref(ms)), "xg_event_elapsed")
            self.total_ms += ms.value
            self.lib.xg_event_destroy(a)
            self.lib.xg_event_destroy(b)
        self.pending = self.pending[n:]


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    fields: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def _summary(results: list, wall: float, busy: float) -> dict:
    lat = np.array([r.latency for r in results])
    wait = np.array([r.admitted - r.arrival for r in results])
    return {"requests": len(results), "wall_s": wall, "solves_per_s": len(results) / wall if wall else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()), "queue_wait_p50_s": float(np.percentile(wait, 50)),
            "gpu_busy_frac": busy / wall if wall else 0.0}


def cells(spec: InstanceSpec) -> int:
    """Cells one request advances per step (summed over its grids)."""
    return (4 if spec.kind == "cavity" else 1) * spec.n * spec.n


class Server:
    def __init__(self, precision: str = "float", hbm_budget: int = 16 << 30, steps_per_tick: int = 8,
                 graphs: bool = False, keep_fields: bool = True, solver=None, timer=None,
                 max_cells: int | None = 64_000_000):
        self.ms = solver if solver is not None else MultiSolver(precision, graphs=graphs)
        self.elem = 4 if precision == "float" else 8
        self.budget = int(hbm_budget)
        # resident-cell cap (None: the HBM budget alone).  Measured on the 200/s FP32
        # cavity trace: 64M cells keeps the throughput of an uncapped batch (6.97
        # against 6.81 solves/s over 400 requests) and halves the median latency
        # (22 s against 44 s); 16M and 8M lose 15-25 % of the throughput.
        self.max_cells = max_cells
        self.steps_per_tick = int(steps_per_tick)
        self.keep_fields = keep_fields
        self.timer = timer if timer is not None else _DeviceTimer()
        self.peak_used = 0
        self.peak_active = 0
        self.ticks = 0

    def run(self, trace: list, time_scale: float = 1.0, on_done=None) -> tuple:
        """Serve `trace` (arrival times in seconds, scaled by `time_scale`) in real time;
        returns (results by request id, summary).  `on_done(result, n_done)` is called
        as each request completes.

        A request whose steps are done leaves the batch at once, and its fields are
        read in the next iteration, after the next tick is queued: the copy waits
        only for the tick it finished in, and the device runs the next one
        meanwhile.  Its memory is released behind the queued kernels, so neither
        admitting nor retiring a request waits for the device."""
        pending = deque(sorted(trace, key=lambda s: s.arrival))
        queue: deque = deque()
        active = {}                      # handle -> [spec, steps done, admitted time]
        retiring = []                    # [spec, admitted time, grids, event after its last tick]
        used = 0
        resident = 0                     # cells of the requests in the batch
        results = {}
        for s in trace:
            if footprint(s, self.elem) > self.budget:
                raise ValueError(f"request {s.rid} needs more than the whole HBM budget")
        self.peak_used = self.peak_active = self.ticks = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0        # noqa: E731
        while pending or queue or active or retiring:
            t = now()
            while pending and pending[0].arrival * time_scale <= t:
                queue.append(pending.popleft())
            while queue and used + footprint(queue[0], self.elem) <= self.budget and \
                    (self.max_cells is None or not active or resident + cells(queue[0]) <= self.max_cells):
                s = queue.popleft()
                h = self.ms.add(s)
                used += footprint(s, self.elem)
                resident += cells(s)
                active[h] = [s, 0, now()]
                self.peak_used = max(self.peak_used, used)
                self.peak_active = max(self.peak_active, len(active))
            if not active and not retiring:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival * time_scale - now()))
                continue
            k = max(0, min(self.steps_per_tick, min((s.steps - d for s, d, _ in active.values()), default=0)))
            if k:
                # no wait here: the host prepares the next tick while the device runs this one
                self.timer.begin()
                self.ms.step(k)
                self.timer.end()
                self.ticks += 1
            event = self.timer.mark()
            for s, adm, grids, after in retiring:
                fields = {}
                if self.keep_fields:           # the copy waits for the request's last tick only
                    names = "bpuv" if s.kind == "cavity" else "u"
                    fields = {c: g._now_after(after).copy() for c, g in zip(names, grids)}
                else:
                    self.timer.wait(after)
                del grids
                used -= footprint(s, self.elem)
                results[s.rid] = Result(s.rid, s.arrival * time_scale, adm, now(), fields)
                if on_done is not None:
                    on_done(results[s.rid], len(results))
            for after in {r[3] for r in retiring}:      # requests of one tick share its event
                self.timer.release(after)
            retiring = []
            finished = []
            for h, rec in active.items():
                rec[1] += k
                if rec[1] >= rec[0].steps:
                    finished.append(h)
            for h in finished:
                s, _, adm = active.pop(h)
                retiring.append([s, adm, self.ms.remove(h), event])
                resident -= cells(s)
            if not finished:
                self.timer.release(event)
        self.timer.sync()
        wall = now()
        busy = self.timer.busy_seconds()
        return results, _summary(list(results.values()), wall, busy)

    def close(self) -> None:
        self.ms.close()


def serve_plain(trace: list, precision: str = "float", time_scale: float = 1.0, keep_fields: bool = True,
                solver: MultiSolver | None = None, on_done=None) -> tuple:
    """The same trace served one request at a time with plain kernel calls (FIFO);
    `on_done(result, n_done)` is called as each request completes."""
    ms = solver or MultiSolver(precision)
    lib = _lib.load()
    results = {}
    busy = 0.0
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0            # noqa: E731
    for s in sorted(trace, key=lambda s: s.arrival):
        wait = s.arrival * time_scale - now()
        if wait > 0:
            time.sleep(wait)
        adm = now()
        grids, scalars = ms.grids_for(s)
        for _ in range(s.steps):
            ms.plain_call(s, grids, scalars)
        _lib.check(lib.xg_synchronize())
        busy += now() - adm
        fields = {}
        if keep_fields:
            names = "bpuv" if s.kind == "cavity" else "u"
            fields = {c: g.now.copy() for c, g in zip(names, grids)}
        del grids
        results[s.rid] = Result(s.rid, s.arrival * time_scale, adm, now(), fields)
        if on_done is not None:
            on_done(results[s.rid], len(results))
    wall = now()
    if solver is None:
        ms.close()
    return results, _summary(list(results.values()), wall, busy)
