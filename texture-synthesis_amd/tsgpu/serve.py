"""Continuous-batching request loop over one device.

Requests (unbuilt SessionBuilders) wait in arrival order. A request is
admitted when the device bytes of the admitted requests plus its own
estimate fit the memory budget and an engine slot is free: it is then built
(pyramids and state go to the device and stay there) and handed to a
``Batch``, whose rounds run every admitted request's planned work in one
launch of the work-queue engine. Requests that arrive while others run join
at the next round, and a finished request frees its memory and slot for the
next in line. Unless ALONE_BELOW or more requests are in flight, a request
starting runs alone with the whole device instead, which shortens its chain
of small batches; the engine takes over as requests pile up. A request whose estimate alone exceeds the budget is rejected
(AdmissionRejected) rather than held; a request can be cancelled; a request
that fails (invalid, out of memory) is reported without disturbing others.

Each result is the same as running the request alone (``Session.run``):
the engine only interleaves launches, it never changes a request's
schedule. The reference has no serving loop; its ``Session::run``
(session.rs:37-66) runs one synthesis per call.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional

from . import Batch, Context, GeneratedImage, SessionBuilder, TsgError

ADMISSION_REJECTED = -24
CANCELLED = -23
# resolve items per request per engine round (0: no bound)
QUANTUM = 0
# a request that starts while fewer than this many are in flight runs alone
# (on the 128^2..2048^2 serving mix: p99 215 against 219 ms at 18 requests/s,
# 320 against 535 ms at 30/s, with 32 rather than 8; the engine pays off only
# once requests pile up)
ALONE_BELOW = 32


@dataclass
class Finished:
    rid: int
    result: Optional[GeneratedImage]
    error: Optional[Exception]
    arrival: float
    admitted: float
    done: float
    footprint: int
    out_px: int

    @property
    def latency_s(self) -> float:
        return self.done - self.arrival


@dataclass
class _Waiting:
    rid: int
    builder: SessionBuilder
    arrival: float
    estimate: int
    out_px: int
    schedule: Optional[dict] = None


@dataclass
class _Running:
    rid: int
    arrival: float
    admitted: float
    footprint: int
    out_px: int


class Server:
    """Admits requests by device-memory footprint and runs them together.

    hbm_budget: device bytes the admitted requests may hold at once.
    max_active: engine slots (requests advanced per round).
    rounds_per_pump: engine rounds run by each ``pump`` call before finished
    requests are collected and waiting ones admitted.
    quantum: resolve items a request may plan per round; it bounds a round's
    length, and so how long a new request waits for the next one.
    alone_below: a request that starts while fewer than this many are in
    flight runs alone, with the whole device; under load requests share the
    work-queue engine.
    """

    def __init__(self, ctx: Optional[Context] = None, hbm_budget: int = 64 << 30, max_active: int = 256,
                 rounds_per_pump: int = 1, quantum: int = QUANTUM, alone_below: int = ALONE_BELOW,
                 clock: Callable[[], float] = time.perf_counter):
        self.ctx = ctx or Context.default()
        self.batch = Batch(self.ctx, max_active=max_active, quantum=quantum, alone_below=alone_below)
        self.alone_below = int(alone_below)
        self.quantum = int(quantum)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.rounds_per_pump = int(rounds_per_pump)
        self.clock = clock
        self.waiting: Deque[_Waiting] = deque()
        self.running: Dict[int, _Running] = {}
        self.in_use = 0
        self.peak_in_use = 0
        self.memory_holds = 0  # admissions deferred because the budget was full
        self.rejected = 0      # requests larger than the whole budget
        self._cancelled: List[Finished] = []
        self._next = 0

    def submit(self, builder: SessionBuilder, arrival: Optional[float] = None, schedule: Optional[dict] = None) -> int:
        """Queues a request; returns its id. arrival defaults to now."""
        rid = self._next
        self._next += 1
        out = builder._out_dims()
        self.waiting.append(_Waiting(rid, builder, self.clock() if arrival is None else arrival,
                                     builder.footprint_estimate(), out.width * out.height, schedule))
        return rid

    def cancel(self, rid: int) -> bool:
        """Cancels a request that has not finished: a waiting one is dropped,
        a running one stops at the engine's next launch. It is reported by
        ``pump`` with a Cancelled error. False when rid is unknown or done."""
        for w in list(self.waiting):
            if w.rid == rid:
                self.waiting.remove(w)
                t = self.clock()
                self._cancelled.append(Finished(rid, None, TsgError(CANCELLED, "request cancelled while waiting"),
                                                w.arrival, t, t, 0, w.out_px))
                return True
        for ticket, r in self.running.items():
            if r.rid == rid:
                self.batch.cancel(ticket)
                return True
        return False

    def _admit(self) -> List[Finished]:
        failed, self._cancelled = self._cancelled, []
        while self.waiting and len(self.running) < self.max_active:
            w = self.waiting[0]
            if w.estimate > self.hbm_budget:
                # can never fit, whatever else finishes: rejected, not held
                self.waiting.popleft()
                t = self.clock()
                err = TsgError(ADMISSION_REJECTED, f"request needs {w.estimate} device bytes, the budget is "
                                                   f"{self.hbm_budget}")
                self.rejected += 1
                failed.append(Finished(w.rid, None, err, w.arrival, t, t, 0, w.out_px))
                continue
            if self.running and self.in_use + w.estimate > self.hbm_budget:
                self.memory_holds += 1
                break
            self.waiting.popleft()
            t = self.clock()
            try:
                w.builder.context(self.ctx)
                sess = w.builder.build()
                if w.schedule:
                    sess.set_schedule(**w.schedule)
                fp = max(w.estimate, sess.footprint())
                ticket = self.batch.submit(sess)
            except Exception as e:  # invalid request: reported, not fatal
                failed.append(Finished(w.rid, None, e, w.arrival, t, self.clock(), 0, w.out_px))
                continue
            self.running[ticket] = _Running(w.rid, w.arrival, t, fp, w.out_px)
            self.in_use += fp
            self.peak_in_use = max(self.peak_in_use, self.in_use)
        return failed

    def pump(self, rounds: Optional[int] = None) -> List[Finished]:
        """Admits what fits, runs engine rounds, returns the requests that finished."""
        done = self._admit()
        if self.running:
            # a request running alone may use the budget's free bytes for its
            # stage-ahead buffers (released when its run ends)
            self.batch.set_alone(self.alone_below, self.hbm_budget - self.in_use)
            self.batch.step(self.rounds_per_pump if rounds is None else rounds)
        for ticket, status in self.batch.poll():
            r = self.running.pop(ticket)
            res, err = None, None
            try:
                res = self.batch.take(ticket)
            except Exception as e:
                err = e
            self.in_use -= r.footprint
            done.append(Finished(r.rid, res, err, r.arrival, r.admitted, self.clock(), r.footprint, r.out_px))
        return done

    def idle(self) -> bool:
        return not self.waiting and not self.running and not self._cancelled

    def drain(self) -> List[Finished]:
        out = []
        while not self.idle():
            out.extend(self.pump())
        return out

    def close(self):
        self.batch.close()


def percentile(values: List[float], q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    if not values:
        return float("nan")
    s = sorted(values)
    i = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[i]
