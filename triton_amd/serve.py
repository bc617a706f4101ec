"""Serving loop: flood-scenario requests stepped together on one GPU.

A thin wrapper over the C++ request loop (tr_serve_*, triton_amd/csrc/host/
serve.cpp): requests (a triton_amd.Case each, run to its sim_duration in the
FP32 form) are admitted into one tr_batch while their device bytes fit the
HBM budget; admitted requests keep their state on the device and advance
together, one batched launch per step for all of them, each with its own dt.
Between chunks of launches the loop retires finished requests (final status
and state kept on the host, device memory freed) and admits waiting ones in
arrival order, so new requests join while others are mid-run (continuous
batching).

A member's arithmetic does not depend on which other members share its
launches, so a request's result is bit-for-bit the result it has alone.

The reference has no serving layer; it runs ensemble members as separate MPI
jobs (src/Ensify.h:43-72).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import Case, SimStatus, check, lib

WAITING, RUNNING, FINISHED, FAILED, CANCELLED = 0, 1, 2, 3, 4


class ServeOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("chunk_steps", C.c_int32), ("waves_per_simd", C.c_int32),
                ("max_members", C.c_int32), ("starve_rounds", C.c_int32), ("keep_state", C.c_int32),
                ("reserved", C.c_int32)]


class ServeRequest(C.Structure):
    _fields_ = [("state", C.c_int32), ("member", C.c_int32), ("bytes", C.c_int64), ("cells", C.c_int64),
                ("skipped", C.c_int32), ("error", C.c_int32), ("arrival", C.c_double),
                ("admitted", C.c_double), ("done", C.c_double), ("status", SimStatus)]


class ServeRequestOptions(C.Structure):
    _fields_ = [("output_dir", C.c_char_p), ("output_format", C.c_int32), ("keep_prints", C.c_int32)]


class ServeStatsC(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("waiting", C.c_int64),
                ("running", C.c_int64), ("cell_steps", C.c_int64), ("rounds", C.c_int64),
                ("peak_bytes", C.c_int64), ("max_members", C.c_int32), ("reserved", C.c_int32),
                ("busy_s", C.c_double), ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double),
                ("max_latency_s", C.c_double), ("failed", C.c_int64), ("cancelled", C.c_int64)]


def _serve_lib() -> C.CDLL:
    L = lib()
    if not getattr(L, "_serve_sig", False):
        vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        from . import CaseView
        sig = {
            "tr_serve_clock": (f64, []),
            "tr_serve_create": (C.c_int, [C.POINTER(ServeOptions), C.POINTER(vp)]),
            "tr_serve_submit": (C.c_int, [vp, C.POINTER(CaseView), f64, C.POINTER(i32)]),
            "tr_serve_step": (C.c_int, [vp, C.POINTER(i32)]),
            "tr_serve_drain": (C.c_int, [vp]),
            "tr_serve_poll": (C.c_int, [vp, i32, C.POINTER(ServeRequest)]),
            "tr_serve_result": (C.c_int, [vp, i32] + [C.POINTER(C.c_double)] * 3),
            "tr_serve_release": (C.c_int, [vp, i32]),
            "tr_serve_submit_ex": (C.c_int, [vp, C.POINTER(CaseView), f64, C.POINTER(ServeRequestOptions),
                                             C.POINTER(i32)]),
            "tr_serve_print_count": (C.c_int, [vp, i32, C.POINTER(i32)]),
            "tr_serve_print": (C.c_int, [vp, i32, i32, C.POINTER(i32), C.POINTER(f64)]
                               + [C.POINTER(C.c_double)] * 4),
            "tr_serve_cancel": (C.c_int, [vp, i32]),
            "tr_serve_stats_get": (C.c_int, [vp, C.POINTER(ServeStatsC)]),
            "tr_serve_destroy": (None, [vp]),
            "tr_serve_admit": (C.c_int, [C.POINTER(i64), C.POINTER(i32), i32, i64, i64, i32, i32, i32,
                                         C.POINTER(C.c_uint8)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._serve_sig = True
    return L


def clock() -> float:
    """The server's clock (seconds): arrival, admitted and done times."""
    return float(_serve_lib().tr_serve_clock())


def admit_round(bytes_, skipped, held: int, budget: int, running: int, max_members: int = 256,
                starve_rounds: int = 8):
    """The admission policy of one round (tr_serve_admit, host only): which
    waiting requests, in arrival order, are admitted; `skipped` (numpy int32)
    is updated in place."""
    b = np.ascontiguousarray(bytes_, dtype=np.int64)
    assert isinstance(skipped, np.ndarray) and skipped.dtype == np.int32 and skipped.flags.c_contiguous
    take = np.zeros(b.shape[0], np.uint8)
    check(_serve_lib().tr_serve_admit(b.ctypes.data_as(C.POINTER(C.c_int64)),
                                      skipped.ctypes.data_as(C.POINTER(C.c_int32)), int(b.shape[0]),
                                      int(held), int(budget), int(running), int(max_members),
                                      int(starve_rounds), take.ctypes.data_as(C.POINTER(C.c_uint8))))
    return take.astype(bool)


@dataclass
class Request:
    rid: int
    arrival: float
    admitted: float | None
    done: float | None
    bytes: int
    cells: int
    steps: int
    skipped: int
    status: dict | None = None
    state: dict | None = None
    tag: object = None
    prints: list = field(default_factory=list)  # (print_id, simtime, state) kept at print times

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
    """Continuous-batching request loop over one tr_batch (tr_serve_*).

    hbm_budget_bytes: device bytes the admitted requests may hold together
        (0: no limit).  A request larger than the whole budget is rejected at
        submit.
    chunk_steps: batched launches between scheduling rounds (admission and
        retirement happen between chunks).
    keep_state: keep the final h, qx, qy of each finished request.
    starve_rounds: a waiting request passed over more than this many times by
        smaller later ones holds back everything behind it until it fits.
    """

    def __init__(self, hbm_budget_bytes: int = 0, chunk_steps: int = 32, keep_state: bool = True,
                 waves_per_simd: int = 0, max_members: int = 256, starve_rounds: int = 8):
        opt = ServeOptions(int(hbm_budget_bytes), int(chunk_steps), int(waves_per_simd), int(max_members),
                           int(starve_rounds), int(bool(keep_state)), 0)
        L = _serve_lib()
        h = C.c_void_p()
        check(L.tr_serve_create(C.byref(opt), C.byref(h)))
        self._h = h
        self.budget = int(hbm_budget_bytes)
        self.keep_state = keep_state
        self._cases: dict[int, Case] = {}     # requests not yet retired
        self._views: dict[int, object] = {}   # their views stay valid until admission
        self._tags: dict[int, object] = {}
        self._done: dict[int, Request] = {}
        self._ended: dict[int, tuple] = {}    # failed or cancelled: (state, TR_* code, tag)
        self._latencies: list[float] = []

    def close(self) -> None:
        if getattr(self, "_h", None):
            _serve_lib().tr_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- client side -------------------------------------------------------
    def submit(self, case: Case, tag=None, arrival: float | None = None, output_dir=None,
               output_format: str = "ASC", keep_prints: bool = False) -> int:
        """Queue a request; `arrival` is on the server's clock (default now).
        output_dir: write the reference's print rasters there at every print
        time (output_format "ASC": ASCII and binary twins, "BIN": binary);
        keep_prints: keep each print's state (Request.prints)."""
        v = case.view()
        rid = C.c_int32()
        ro = ServeRequestOptions(None if output_dir is None else str(output_dir).encode(),
                                 0 if str(output_format).upper() == "ASC" else 1, int(bool(keep_prints)))
        check(_serve_lib().tr_serve_submit_ex(self._h, C.byref(v), -1.0 if arrival is None else float(arrival),
                                              C.byref(ro), C.byref(rid)))
        r = rid.value
        self._cases[r], self._views[r], self._tags[r] = case, v, tag
        return r

    def poll(self, rid: int) -> ServeRequest:
        out = ServeRequest()
        check(_serve_lib().tr_serve_poll(self._h, int(rid), C.byref(out)))
        return out

    def _stats(self) -> ServeStatsC:
        st = ServeStatsC()
        check(_serve_lib().tr_serve_stats_get(self._h, C.byref(st)))
        return st

    def pending(self) -> int:
        st = self._stats()
        return int(st.waiting + st.running)

    @property
    def running(self) -> list[int]:
        return [r for r in self._cases if self.poll(r).state == RUNNING]

    @property
    def waiting(self) -> list[int]:
        return [r for r in self._cases if self.poll(r).state == WAITING]

    # -- scheduler ----------------------------------------------------------
    def _collect(self) -> None:
        """Move finished requests' results to the host side and release them."""
        L = _serve_lib()
        for r in list(self._cases):
            p = self.poll(r)
            if p.state == RUNNING:
                self._views.pop(r, None)
            if p.state in (FAILED, CANCELLED):
                self._cases.pop(r)
                self._views.pop(r, None)
                self._ended[r] = (p.state, int(p.error), self._tags.pop(r, None))
                check(L.tr_serve_release(self._h, int(r)))
                continue
            if p.state != FINISHED:
                continue
            case = self._cases.pop(r)
            self._views.pop(r, None)
            state = None
            if self.keep_state:
                shape = (case.rows, case.cols)
                state = {k: np.empty(shape, np.float64) for k in ("h", "qx", "qy")}
                P = C.POINTER(C.c_double)
                check(L.tr_serve_result(self._h, int(r), *(state[k].ctypes.data_as(P) for k in ("h", "qx", "qy"))))
            status = p.status.as_dict()
            prints = []
            n = C.c_int32()
            check(L.tr_serve_print_count(self._h, int(r), C.byref(n)))
            P = C.POINTER(C.c_double)
            for k in range(n.value):
                shape = (case.rows, case.cols)
                ps = {f: np.empty(shape, np.float64) for f in ("h", "qx", "qy", "maxh")}
                pid, t = C.c_int32(), C.c_double()
                check(L.tr_serve_print(self._h, int(r), k, C.byref(pid), C.byref(t),
                                       *(ps[f].ctypes.data_as(P) for f in ("h", "qx", "qy", "maxh"))))
                if "maxh" not in case.arrays:
                    ps.pop("maxh")
                prints.append((pid.value, t.value, ps))
            self._done[r] = Request(r, p.arrival, p.admitted, p.done, int(p.bytes), int(p.cells),
                                    int(status["steps"]), int(p.skipped), status, state, self._tags.pop(r, None),
                                    prints)
            self._latencies.append(p.done - p.arrival)
            check(L.tr_serve_release(self._h, int(r)))

    def step(self) -> bool:
        """One scheduling round: admit, run one chunk, retire.  Returns False
        when nothing was on the device."""
        ran = C.c_int32()
        check(_serve_lib().tr_serve_step(self._h, C.byref(ran)))
        self._collect()
        return bool(ran.value)

    def drain(self) -> None:
        """Run until every submitted request has finished."""
        while self.pending():
            if not self.step() and self.waiting:
                raise RuntimeError("waiting requests cannot be admitted")

    def cancel(self, rid: int) -> None:
        """End a waiting or running request now (tr_serve_cancel)."""
        check(_serve_lib().tr_serve_cancel(self._h, int(rid)))
        self._collect()

    def ended(self, rid: int):
        """(state, TR_* code, tag) of a failed or cancelled request, or None."""
        return self._ended.get(rid)

    def result(self, rid: int) -> Request:
        return self._done[rid]

    @property
    def stats(self) -> ServeStats:
        st = self._stats()
        return ServeStats(int(st.completed), int(st.cell_steps), float(st.busy_s), int(st.rounds),
                          int(st.max_members), int(st.peak_bytes), list(self._latencies))


def serve_poisson(requests, rate_per_s: float, seed: int = 5, hbm_budget_bytes: int = 0,
                  chunk_steps: int = 32, keep_state: bool = False, log=None, waves_per_simd: int = 0):
    """Replay `requests` (Cases) arriving as a Poisson process of the given
    rate against a Server; returns (server, arrival offsets).  Arrivals are
    checked between chunks, so a request waits at most one chunk before it
    can be admitted."""
    rng = np.random.default_rng(seed)
    offsets = np.cumsum(rng.exponential(1.0 / rate_per_s, size=len(requests)))
    offsets -= offsets[0]
    srv = Server(hbm_budget_bytes, chunk_steps, keep_state=keep_state, waves_per_simd=waves_per_simd)
    t0 = clock()
    k = 0
    while k < len(requests) or srv.pending():
        now = clock() - t0
        while k < len(requests) and offsets[k] <= now:
            srv.submit(requests[k], tag=k, arrival=t0 + offsets[k])
            k += 1
            if log:
                log(f"t={now:7.2f}s submit {k - 1} ({requests[k - 1].rows - 2}^2), "
                    f"pending {srv.pending()}")
        if not srv.step() and k < len(requests):
            time.sleep(max(0.0, min(0.05, offsets[k] - (clock() - t0))))
    return srv, offsets
