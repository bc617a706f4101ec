"""Serving loop: flood-scenario requests stepped together on one GPU.

Requests (a triton_amd.Case each, run to its sim_duration in the FP32 form)
are admitted into one tr_batch while their device bytes fit the HBM budget;
admitted requests keep their state on the device and advance together, one
batched launch per step for all of them, each with its own dt.  Between
chunks of launches the loop retires finished requests (downloads their final
state, frees their memory) and admits waiting ones in arrival order, so new
requests join while others are mid-run (continuous batching).

A member's arithmetic does not depend on which other members share its
launches, so a request's result is bit-for-bit the result it has alone.

The reference has no serving layer; it runs ensemble members as separate MPI
jobs (src/Ensify.h:43-72).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Case


@dataclass
class Request:
    rid: int
    case: Case
    arrival: float                  # host clock (time.perf_counter) at submission
    bytes: int
    admitted: float | None = None
    done: float | None = None
    member: int | None = None
    status: dict | None = None
    state: dict | None = None
    cells: int = 0
    steps: int = 0
    skipped: int = 0                # admission rounds it was passed over
    tag: object = None

    @property
    def latency(self) -> float | None:
        return None if self.done is None else self.done - self.arrival


@dataclass
class ServeStats:
    completed: int = 0
    cell_steps: int = 0             # sum over retired requests of interior cells x steps
    busy_s: float = 0.0             # wall time with at least one request on the device
    rounds: int = 0
    max_members: int = 0
    peak_bytes: int = 0
    latencies: list = field(default_factory=list)

    def percentile(self, q: float) -> float | None:
        return float(np.percentile(self.latencies, q)) if self.latencies else None


class Server:
    """Continuous-batching request loop over one tr_batch.

    hbm_budget_bytes: device bytes the admitted requests may hold together
        (0: no limit).  A request larger than the whole budget is rejected at
        submit.
    chunk_steps: batched launches between scheduling rounds (admission and
        retirement happen between chunks).
    keep_state: download the final h, qx, qy of each finished request.
    starve_rounds: a waiting request passed over this many times by smaller
        later ones holds back everything behind it until it fits.
    batch: the tr_batch wrapper to drive (default: a new Batch with the
        budget); anything with the Batch methods used here will do.
    """

    def __init__(self, hbm_budget_bytes: int = 0, chunk_steps: int = 32, keep_state: bool = True,
                 waves_per_simd: int = 0, max_members: int = 256, starve_rounds: int = 8, batch=None):
        if chunk_steps < 1:
            raise ValueError("chunk_steps must be >= 1")
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk_steps)
        self.keep_state = keep_state
        self.max_members = int(max_members)
        self.starve_rounds = int(starve_rounds)
        self.batch = batch if batch is not None else Batch(hbm_budget_bytes=self.budget,
                                                            waves_per_simd=waves_per_simd)
        self.waiting: list[Request] = []
        self.running: dict[int, Request] = {}
        self.finished: dict[int, Request] = {}
        self.stats = ServeStats()
        self._next = 0

    def close(self) -> None:
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- client side -------------------------------------------------------
    def submit(self, case: Case, tag=None, arrival: float | None = None) -> int:
        need = self.batch.member_bytes(case)
        if need < 0:
            raise ValueError("case cannot run in the batched FP32 form")
        if self.budget > 0 and need > self.budget:
            raise ValueError(f"request needs {need} B, more than the whole HBM budget {self.budget} B")
        rid = self._next
        self._next += 1
        self.waiting.append(Request(rid, case, time.perf_counter() if arrival is None else arrival, need,
                                    cells=(case.rows - 2) * (case.cols - 2), tag=tag))
        return rid

    def pending(self) -> int:
        return len(self.waiting) + len(self.running)

    # -- scheduler ----------------------------------------------------------
    def _admit(self) -> None:
        held = self.batch.device_bytes
        keep = []
        blocked = False
        for req in self.waiting:  # arrival order, first fit
            fits = (self.budget <= 0 or held + req.bytes <= self.budget) and \
                len(self.running) < self.max_members
            if blocked or not fits:
                if not fits:
                    req.skipped += 1
                    blocked = blocked or req.skipped > self.starve_rounds
                keep.append(req)
                continue
            req.member = self.batch.add(req.case)
            req.admitted = time.perf_counter()
            held = self.batch.device_bytes
            self.running[req.member] = req
        self.waiting = keep
        self.stats.max_members = max(self.stats.max_members, len(self.running))
        self.stats.peak_bytes = max(self.stats.peak_bytes, held)

    def _retire(self) -> None:
        for mid in list(self.running):
            st = self.batch.poll(mid)
            if not st["finished"]:
                continue
            req = self.running.pop(mid)
            req.status = self.batch.status(mid)
            if self.keep_state:
                req.state = self.batch.state(mid)
            req.steps = int(req.status["steps"])
            req.done = time.perf_counter()
            self.batch.remove(mid)
            req.member = None
            self.finished[req.rid] = req
            self.stats.completed += 1
            self.stats.cell_steps += req.cells * req.steps
            self.stats.latencies.append(req.latency)

    def step(self) -> bool:
        """One scheduling round: admit, run one chunk, retire.  Returns False
        when nothing was on the device."""
        self._admit()
        if not self.running:
            return False
        t0 = time.perf_counter()
        self.batch.arm(max_steps=self.chunk)
        self.batch.enqueue(self.chunk)
        self.batch.sync()
        self._retire()
        self.stats.busy_s += time.perf_counter() - t0
        self.stats.rounds += 1
        return True

    def drain(self) -> None:
        """Run until every submitted request has finished."""
        while self.pending():
            if not self.step() and self.waiting:
                raise RuntimeError("waiting requests cannot be admitted")

    def result(self, rid: int) -> Request:
        return self.finished[rid]


def serve_poisson(requests, rate_per_s: float, seed: int = 5, hbm_budget_bytes: int = 0,
                  chunk_steps: int = 32, keep_state: bool = False, log=None):
    """Replay `requests` (Cases) arriving as a Poisson process of the given
    rate against a Server; returns (server, arrival offsets).  Arrivals are
    checked between chunks, so a request waits at most one chunk before it
    can be admitted."""
    rng = np.random.default_rng(seed)
    offsets = np.cumsum(rng.exponential(1.0 / rate_per_s, size=len(requests)))
    offsets -= offsets[0]
    srv = Server(hbm_budget_bytes, chunk_steps, keep_state=keep_state)
    t0 = time.perf_counter()
    k = 0
    while k < len(requests) or srv.pending():
        now = time.perf_counter() - t0
        while k < len(requests) and offsets[k] <= now:
            srv.submit(requests[k], tag=k, arrival=t0 + offsets[k])
            k += 1
            if log:
                log(f"t={now:7.2f}s submit {k - 1} ({requests[k - 1].rows - 2}^2), "
                    f"running {len(srv.running)}, waiting {len(srv.waiting)}")
        if not srv.step() and k < len(requests):
            time.sleep(max(0.0, min(0.05, offsets[k] - (time.perf_counter() - t0))))
    return srv, offsets
