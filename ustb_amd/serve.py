"""Serving loop: continuous batching of beamforming requests on one GPU.

Requests (one frame each, of one of several probe configurations) arrive on
a trace. The server keeps every configuration's plan (geometry, delays and
apodization tables) resident, admits a request when its input and output
buffers fit in the HBM budget (smallest input first, with a bounded amount
of input and at most one large frame queued on the copy stream), copies its
channel data in on a copy stream,
and runs iterations of the batcher on a compute stream: each iteration takes
the ready requests in order of least remaining work (then arrival) and gives
each a slice of its tiles until about `slice_ms` of estimated work is packed
(`long_slices` times that while no small request is pending; small frames go
whole, large frames over several iterations), so requests
that arrive while others are on the GPU join the next iteration and small
frames are not held behind large ones. Up to `depth` iterations are in
flight. A request whose last slice has run is copied back on a third stream;
its latency is arrival -> result in host memory.

`plain_loop` is the baseline: one request at a time, one plan launch per
frame (copy in, beamform, copy out, wait).

Every request ends with a status (ServeResult.status):
- "ok": its image is in host memory (latency measured);
- "rejected": its input and output alone exceed the HBM budget, so it is
  never admitted;
- "failed": the library refused its work (a C-ABI error on submission). A
  submission that fails is retried one request at a time, so the error stays
  with the request that caused it and the others of the iteration complete;
- "cancelled": run()'s cancel callback returned true for it, or it was not
  done within its Request.deadline_s of its arrival. Work already queued for
  it finishes on the GPU before its buffers are released; nothing more is
  queued for it.
A fault of the device itself (a kernel that faults) ends the process's GPU
context and cannot be confined to one request.

roctx marks (ustb_amd.trace) name each admission, iteration and completion
for rocprofv3 --marker-trace.
"""
from __future__ import annotations

import ctypes
import heapq
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import trace as tr
from .ops import GeometryPlan, _stream_ptr


@dataclass
class Request:
    rid: int
    config: int
    arrival: float          # seconds after the trace starts
    data: int               # dataset index of its configuration
    deadline_s: float | None = None  # cancelled if not done this long after arrival


def poisson_trace(n: int, rate: float, n_configs: int, n_data: int, seed: int = 5) -> list[Request]:
    """n requests with exponential inter-arrival times (mean 1/rate), each
    drawing a configuration uniformly and one of n_data datasets."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    t -= t[0]
    cfg = rng.integers(0, n_configs, n)
    dat = rng.integers(0, n_data, n)
    return [Request(i, int(c), float(a), int(d)) for i, (a, c, d) in enumerate(zip(t, cfg, dat))]


class Batcher:
    """ctypes wrapper of ustbx_batcher."""

    def __init__(self, interp: int, max_acqs: int, max_items: int, wave_chunk: int, split: int):
        self.handle = ctypes.c_void_p()
        st = L.Status()
        L.check(L.lib().ustbx_batcher_create(interp, max_acqs, max_items, wave_chunk, split,
                                             ctypes.byref(self.handle), ctypes.byref(st)), st)

    def submit(self, plans, chs, outs, begin, end, stream):
        n = len(plans)
        P = (ctypes.c_void_p * n)(*[p.handle.value for p in plans])
        C = (ctypes.c_void_p * n)(*[c.data_ptr() for c in chs])
        O = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        B = (ctypes.c_int32 * n)(*begin)
        E = (ctypes.c_int32 * n)(*end)
        st = L.Status()
        L.check(L.lib().ustbx_batcher_submit(self.handle, n, P, C, O, B, E, stream.cuda_stream, ctypes.byref(st)),
                st)

    def close(self):
        if self.handle:
            L.lib().ustbx_batcher_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class _Live:
    req: Request
    bytes: int
    ch: torch.Tensor
    out: torch.Tensor
    ready: torch.cuda.Event
    next_pos: int = 0
    done: torch.cuda.Event | None = None
    host: torch.Tensor | None = None
    last: torch.cuda.Event | None = None  # the last iteration that ran some of its work


@dataclass
class ServeResult:
    latency_s: np.ndarray            # per request, in trace order (NaN unless "ok")
    configs: np.ndarray              # configuration per request
    elapsed_s: float                 # first arrival -> last result
    iterations: int
    outputs: dict = field(default_factory=dict)   # rid -> host image (keep_outputs)
    status: list = field(default_factory=list)    # per request: ok / rejected / failed / cancelled
    errors: dict = field(default_factory=dict)    # rid -> message of a failed or rejected request

    @property
    def ok(self) -> np.ndarray:
        if not self.status:
            return np.ones(len(self.latency_s), bool)
        return np.array([s == "ok" for s in self.status])

    @property
    def frames_per_s(self) -> float:
        return int(self.ok.sum()) / self.elapsed_s

    def summary(self, names: list[str] | None = None) -> dict:
        ok = self.ok
        lat = self.latency_s[ok] * 1e3
        cfg = self.configs[ok]
        res = {"frames": int(lat.size), "frames_per_s": round(self.frames_per_s, 2),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "elapsed_s": round(self.elapsed_s, 3)}
        if self.status:
            res["status"] = {k: int(sum(s == k for s in self.status)) for k in ("ok", "rejected", "failed",
                                                                              "cancelled")}
        if self.iterations:
            res["iterations"] = self.iterations
        per = {}
        for k in np.unique(cfg):
            m = cfg == k
            key = names[k] if names else str(int(k))
            per[key] = {"frames": int(m.sum()), "p50_ms": round(float(np.percentile(lat[m], 50)), 3),
                        "p99_ms": round(float(np.percentile(lat[m], 99)), 3)}
        res["per_config"] = per
        return res


class Server:
    def __init__(self, plans: list[GeometryPlan], datasets: list[list[np.ndarray]], *,
                 hbm_budget: int = 64 << 30, slice_ms: float = 2.0, depth: int = 2,
                 tiles: tuple[int, int] = (16, 4), max_acqs: int = 4096, copy_window: int = 512 << 20,
                 long_slices: int = 8, idle_slice_ms: float | None = 1.0, idle_long_slices: int | None = 4,
                 idle_depth: int | None = 1, idle_below_ms: float = 250.0):
        """slice_ms / long_slices / depth: iteration length (short while a
        small frame is pending, long_slices times that otherwise) and the
        number of iterations queued on the GPU under a backlog (throughput
        first). The idle_* values apply while the work not yet submitted
        (arrived requests' remaining positions at their calibrated cost) is
        below idle_below_ms (latency first: shorter iterations and one queued
        iteration, so a small frame that arrives waits behind less work).
        None keeps the backlog value."""
        if not plans:
            raise ValueError("no configurations")
        interp = plans[0].interp
        if any(p.interp != interp or p.dim != L.DIM_BOTH for p in plans):
            raise ValueError("the configurations share one interpolation and dimension both")
        for p in plans:
            if p.variant[0] < 2:
                p.set_variant(2)
        self.plans = plans
        self.device = plans[0].device
        self.budget = int(hbm_budget)
        self.copy_window = int(copy_window)
        self.long_slices = int(long_slices)
        self.slice_ms = float(slice_ms)
        self.depth = int(depth)
        self.idle_slice_ms = float(slice_ms if idle_slice_ms is None else idle_slice_ms)
        self.idle_long_slices = int(long_slices if idle_long_slices is None else idle_long_slices)
        self.idle_depth = int(depth if idle_depth is None else idle_depth)
        self.idle_below_ms = float(idle_below_ms)
        if min(self.depth, self.idle_depth, self.long_slices, self.idle_long_slices) < 1 or \
                min(self.slice_ms, self.idle_slice_ms) < 0:
            raise ValueError("depth and long slices must be positive, slices not negative")
        self.positions = [int(L.lib().ustbx_plan_positions(p.handle)) for p in plans]
        for k, n in enumerate(self.positions):
            if n < 0:
                raise ValueError(f"configuration {k}: the batcher needs a dimension-both plan with receive-table "
                                 "state")
        # host channel data in pinned memory (the requests' payloads)
        self.host = [[torch.from_numpy(np.ascontiguousarray(d)).pin_memory() for d in ds] for ds in datasets]
        self.in_bytes = [int(ds[0].nbytes) for ds in datasets]
        self.out_bytes = [int(np.prod(p.out_shape)) * 8 for p in plans]
        tiles_per_plan = [((p.P + 63) // 64) * p.F for p in plans]
        self.max_items = max(4 * sum(tiles_per_plan), 1 << 16)
        self.max_acqs = int(max_acqs)
        self.batcher = Batcher(interp, self.max_acqs, self.max_items, *tiles)
        self.tiles = tiles
        self.s_in = torch.cuda.Stream(self.device)
        self.s_run = torch.cuda.Stream(self.device)
        self.s_out = torch.cuda.Stream(self.device)
        self.ms_per_pos = self._calibrate()

    def _small(self, k: int) -> bool:
        """A configuration whose whole frame fits in one short iteration."""
        return self.ms_per_pos[k] * self.positions[k] <= self.slice_ms

    def _calibrate(self) -> list[float]:
        """Time one whole frame of every configuration through the batcher;
        the per-position cost drives slicing."""
        res = []
        with torch.cuda.device(self.device):
            for k, p in enumerate(self.plans):
                with torch.cuda.stream(self.s_run):
                    ch = self.host[k][0].to(self.device, non_blocking=True)
                    out = torch.empty(p.out_shape, dtype=torch.complex64, device=self.device)
                for rep in range(2):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(self.s_run)
                    self.batcher.submit([p], [ch], [out], [0], [self.positions[k]], self.s_run)
                    e1.record(self.s_run)
                    e1.synchronize()
                res.append(e0.elapsed_time(e1) / max(self.positions[k], 1))
                del ch, out
        return res

    def _submit(self, plans, chs, outs, begin, end):
        self.batcher.submit(plans, chs, outs, begin, end, self.s_run)

    def run(self, trace: list[Request], keep_outputs: bool = False, cancel=None) -> ServeResult:
        """Serve the trace; cancel(rid) -> bool is asked about every request
        not yet done on every pass of the loop (a true answer cancels it)."""
        dev = self.device
        n = len(trace)
        lat = np.full(n, np.nan)
        status = ["ok"] * n
        errors = {}
        waiting: list = []  # heap of (input bytes, arrival, rid, request)
        staged: deque[_Live] = deque()
        copying = 0  # input bytes queued on the copy stream
        copying_large = 0  # of them, frames too large for one short iteration
        ready: list[_Live] = []
        running: deque[tuple[torch.cuda.Event, list[_Live]]] = deque()
        draining: deque[_Live] = deque()
        releasing: list[_Live] = []  # cancelled or failed, waiting for their queued work
        outputs = {}
        in_flight = 0
        completed = 0
        nxt = 0
        iters = 0
        t0 = time.perf_counter()
        last_done = 0.0
        backlog_ms = 0.0  # arrived work not yet submitted, at the calibrated cost

        def stop(r: Request, why: str) -> bool:
            """Whether request r is to be cancelled now."""
            if cancel is not None and cancel(r.rid):
                return True
            return r.deadline_s is not None and (time.perf_counter() - t0) - r.arrival > r.deadline_s

        def release(lv: _Live, st: str, msg: str | None = None):
            """End a request that did not complete: its queued work (if any)
            finishes before its buffers go."""
            nonlocal backlog_ms
            status[lv.req.rid] = st
            if msg:
                errors[lv.req.rid] = msg
            backlog_ms -= (self.positions[lv.req.config] - lv.next_pos) * self.ms_per_pos[lv.req.config]
            lv.next_pos = self.positions[lv.req.config]
            releasing.append(lv)
            tr.mark(f"{st} rid={lv.req.rid}")

        with torch.cuda.device(dev):
            while completed < n:
                now = time.perf_counter() - t0
                progressed = False
                while nxt < n and trace[nxt].arrival <= now:
                    r = trace[nxt]
                    nxt += 1
                    need = self.in_bytes[r.config] + self.out_bytes[r.config]
                    if need > self.budget:  # can never be admitted
                        status[r.rid] = "rejected"
                        errors[r.rid] = f"needs {need} bytes of the {self.budget}-byte HBM budget"
                        tr.mark(f"rejected rid={r.rid}")
                        completed += 1
                        continue
                    heapq.heappush(waiting, (self.in_bytes[r.config], r.arrival, r.rid, r))
                    backlog_ms += self.positions[r.config] * self.ms_per_pos[r.config]
                # cancellations: not yet admitted, being copied in, or ready
                if waiting and (cancel is not None or any(w[3].deadline_s is not None for w in waiting)):
                    keep = []
                    for w in waiting:
                        if stop(w[3], "cancelled"):
                            status[w[3].rid] = "cancelled"
                            backlog_ms -= self.positions[w[3].config] * self.ms_per_pos[w[3].config]
                            tr.mark(f"cancelled rid={w[3].rid}")
                            completed += 1
                            progressed = True
                        else:
                            keep.append(w)
                    if len(keep) != len(waiting):
                        waiting = keep
                        heapq.heapify(waiting)
                for lv in list(ready):
                    if stop(lv.req, "cancelled"):
                        ready.remove(lv)
                        release(lv, "cancelled")
                        progressed = True
                # admission: smallest input first, under the HBM budget, with
                # at most copy_window bytes and one large frame's input queued
                # on the copy stream (so a small frame's copy waits behind at
                # most one large copy)
                while waiting:
                    r = waiting[0][3]
                    need = self.in_bytes[r.config] + self.out_bytes[r.config]
                    if in_flight + need > self.budget:
                        break
                    if copying + self.in_bytes[r.config] > self.copy_window and copying > 0:
                        break
                    large = not self._small(r.config)
                    if large and copying_large > 0:
                        break
                    heapq.heappop(waiting)
                    copying += self.in_bytes[r.config]
                    copying_large += 1 if large else 0
                    p = self.plans[r.config]
                    ch = torch.empty((p.F, p.W, p.N, p.L), dtype=torch.complex64, device=dev)
                    out = torch.empty(p.out_shape, dtype=torch.complex64, device=dev)
                    with torch.cuda.stream(self.s_in):
                        ch.copy_(self.host[r.config][r.data], non_blocking=True)
                        ev = torch.cuda.Event()
                        ev.record(self.s_in)
                    staged.append(_Live(r, need, ch, out, ev))
                    in_flight += need
                    tr.mark(f"admit rid={r.rid} config={r.config}")
                    progressed = True
                while staged and staged[0].ready.query():
                    lv = staged.popleft()
                    copying -= self.in_bytes[lv.req.config]
                    copying_large -= 0 if self._small(lv.req.config) else 1
                    if stop(lv.req, "cancelled"):
                        release(lv, "cancelled")
                    else:
                        ready.append(lv)
                    progressed = True
                # an iteration over the ready requests; the settings follow the
                # load: the arrived work not yet submitted
                idle = backlog_ms < self.idle_below_ms
                if ready and len(running) < (self.idle_depth if idle else self.depth):
                    group, begin, end = [], [], []
                    # a short iteration while small requests are pending (their
                    # latency), a long one when only large frames are (fewer,
                    # fuller launches)
                    small = any(self._small(lv.req.config) for lv in ready) or \
                        any(self._small(lv.req.config) for lv in staged) or \
                        (bool(waiting) and self._small(waiting[0][3].config))
                    sl, lg = (self.idle_slice_ms, self.idle_long_slices) if idle else (self.slice_ms, self.long_slices)
                    budget = sl if small else sl * lg
                    items = 0
                    # shortest remaining work first (then arrival): small
                    # frames pass large ones, and the large frame nearest
                    # completion gets the rest of the budget
                    ready.sort(key=lambda lv: ((self.positions[lv.req.config] - lv.next_pos) *
                                               self.ms_per_pos[lv.req.config], lv.req.arrival))
                    for lv in ready:
                        if group and (budget <= 0 or len(group) == self.max_acqs):
                            break  # (the oldest ready request always gets a slice)
                        k = lv.req.config
                        left = self.positions[k] - lv.next_pos
                        take = min(left, max(1, int(budget / max(self.ms_per_pos[k], 1e-9))))
                        per_pos = 8 * self.plans[k].F  # work items per position (at most)
                        if items + take * per_pos > self.max_items:
                            take = (self.max_items - items) // per_pos
                            if take <= 0:
                                break
                        items += take * per_pos
                        group.append(lv)
                        begin.append(lv.next_pos)
                        end.append(lv.next_pos + take)
                        budget -= self.ms_per_pos[k] * take
                    with tr.range_(f"iteration {iters}: {len(group)} requests"):
                        try:
                            self._submit([self.plans[lv.req.config] for lv in group], [lv.ch for lv in group],
                                         [lv.out for lv in group], begin, end)
                            done_ok = list(range(len(group)))
                        except L.USTBXError:
                            # one request at a time: the error stays with its request
                            done_ok = []
                            for i, lv in enumerate(group):
                                try:
                                    self._submit([self.plans[lv.req.config]], [lv.ch], [lv.out], [begin[i]],
                                                 [end[i]])
                                    done_ok.append(i)
                                except L.USTBXError as e:
                                    release(lv, "failed", str(e))
                    ev = torch.cuda.Event()
                    ev.record(self.s_run)
                    finishing = []
                    for i in done_ok:
                        lv = group[i]
                        k = lv.req.config
                        take = end[i] - begin[i]
                        lv.next_pos += take
                        lv.last = ev
                        backlog_ms -= self.ms_per_pos[k] * take
                        if lv.next_pos >= self.positions[k]:
                            finishing.append(lv)
                    running.append((ev, finishing))
                    ready = [lv for lv in ready if lv.next_pos < self.positions[lv.req.config]]
                    iters += 1
                    progressed = True
                # finished iterations: copy their completed requests out
                while running and running[0][0].query():
                    ev, finishing = running.popleft()
                    for lv in finishing:
                        self.s_out.wait_event(ev)
                        with torch.cuda.stream(self.s_out):
                            lv.host = torch.empty(lv.out.shape, dtype=lv.out.dtype, pin_memory=True)
                            lv.host.copy_(lv.out, non_blocking=True)
                            lv.done = torch.cuda.Event()
                            lv.done.record(self.s_out)
                        draining.append(lv)
                    progressed = True
                while draining and draining[0].done.query():
                    lv = draining.popleft()
                    tdone = time.perf_counter() - t0
                    lat[lv.req.rid] = tdone - lv.req.arrival
                    last_done = max(last_done, tdone)
                    if keep_outputs:
                        outputs[lv.req.rid] = lv.host.numpy()
                    in_flight -= lv.bytes
                    lv.ch = lv.out = lv.host = None
                    completed += 1
                    tr.mark(f"complete rid={lv.req.rid}")
                    progressed = True
                # cancelled or failed requests: buffers go once their queued work is done
                for lv in list(releasing):
                    if lv.last is None or lv.last.query():
                        releasing.remove(lv)
                        in_flight -= lv.bytes
                        lv.ch = lv.out = lv.host = None
                        completed += 1
                        progressed = True
                if not progressed:
                    time.sleep(20e-6)
        cfgs = np.array([r.config for r in trace])
        ok = [status[r.rid] == "ok" for r in trace]
        first = min((r.arrival for r, o in zip(trace, ok) if o), default=0.0)
        return ServeResult(lat, cfgs, max(last_done - first, 1e-9), iters, outputs, status, errors)

    def plain_loop(self, trace: list[Request], keep_outputs: bool = False) -> ServeResult:
        """One request at a time: copy in, one launch of its plan, copy out, wait."""
        dev = self.device
        n = len(trace)
        lat = np.zeros(n)
        outputs = {}
        bufs = {}
        s = self.s_run
        t0 = time.perf_counter()
        last_done = 0.0
        with torch.cuda.device(dev):
            for k, p in enumerate(self.plans):
                bufs[k] = (torch.empty((p.F, p.W, p.N, p.L), dtype=torch.complex64, device=dev),
                           torch.empty(p.out_shape, dtype=torch.complex64, device=dev),
                           torch.empty(p.out_shape, dtype=torch.complex64, pin_memory=True))
            for r in trace:
                while time.perf_counter() - t0 < r.arrival:
                    time.sleep(20e-6)
                ch, out, host = bufs[r.config]
                with torch.cuda.stream(s):
                    ch.copy_(self.host[r.config][r.data], non_blocking=True)
                    self.plans[r.config].run(ch, out, stream=s)
                    host.copy_(out, non_blocking=True)
                s.synchronize()
                tdone = time.perf_counter() - t0
                lat[r.rid] = tdone - r.arrival
                last_done = tdone
                if keep_outputs:
                    outputs[r.rid] = host.numpy().copy()
        cfgs = np.array([r.config for r in trace])
        return ServeResult(lat, cfgs, last_done - trace[0].arrival, 0, outputs)

    def close(self):
        self.batcher.close()
