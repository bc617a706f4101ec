"""Scheduling logic of the serving loop on the CPU, against a stand-in for
the batch (no device): admission within the HBM budget, arrival order with
first fit, the starvation guard, retirement, and latency bookkeeping."""
import pytest

from triton_amd.serve import Server


class FakeCase:
    def __init__(self, side, nbytes, steps):
        self.rows = self.cols = side + 2
        self.nbytes = nbytes
        self.need = steps


class FakeBatch:
    """Steps every member by the chunk; a member finishes after its
    `need` steps.  Mirrors the Batch methods the server calls."""

    def __init__(self):
        self.m = {}
        self.next = 0
        self.log = []

    @staticmethod
    def member_bytes(case):
        return case.nbytes

    @property
    def device_bytes(self):
        return sum(c.nbytes for c, _ in self.m.values())

    def add(self, case):
        mid = self.next
        self.next += 1
        self.m[mid] = [case, 0]
        self.log.append(("add", mid))
        return mid

    def remove(self, mid):
        del self.m[mid]
        self.log.append(("remove", mid))

    def arm(self, max_steps):
        self.limit = max_steps

    def enqueue(self, n):
        for rec in self.m.values():
            rec[1] = min(rec[0].need, rec[1] + min(n, self.limit))

    def sync(self):
        return sum(1 for c, s in self.m.values() if s < c.need)

    def poll(self, mid):
        c, s = self.m[mid]
        return {"finished": s >= c.need, "steps": s}

    def status(self, mid):
        return self.poll(mid)

    def state(self, mid):
        return {"steps": self.m[mid][1]}

    def close(self):
        pass


def test_budget_is_never_exceeded_and_all_requests_finish():
    fb = FakeBatch()
    srv = Server(hbm_budget_bytes=100, chunk_steps=10, batch=fb)
    sizes = [60, 30, 50, 10, 40, 70]
    for i, b in enumerate(sizes):
        srv.submit(FakeCase(8 + i, b, 25 + 10 * i))
    peak = 0
    while srv.pending():
        srv.step()
        peak = max(peak, fb.device_bytes)
        assert srv.stats.peak_bytes <= 100
    assert peak <= 100
    assert srv.stats.completed == len(sizes)
    for rid in range(len(sizes)):
        r = srv.result(rid)
        assert r.steps == 25 + 10 * rid
        assert r.latency is not None and r.latency >= 0
        assert r.admitted >= r.arrival and r.done >= r.admitted
    assert srv.stats.cell_steps == sum((8 + i) ** 2 * (25 + 10 * i) for i in range(len(sizes)))


def test_first_fit_in_arrival_order():
    fb = FakeBatch()
    srv = Server(hbm_budget_bytes=100, chunk_steps=5, batch=fb)
    srv.submit(FakeCase(4, 70, 50))   # 0: admitted
    srv.submit(FakeCase(4, 50, 50))   # 1: does not fit next to 0
    srv.submit(FakeCase(4, 20, 50))   # 2: fits, admitted before 1
    srv.step()
    assert [e for e in fb.log if e[0] == "add"] == [("add", 0), ("add", 1)]
    admitted = {r.rid for r in srv.running.values()}
    assert admitted == {0, 2}
    assert [r.rid for r in srv.waiting] == [1]


def test_starved_request_holds_back_later_ones():
    fb = FakeBatch()
    srv = Server(hbm_budget_bytes=100, chunk_steps=1, batch=fb, starve_rounds=2)
    srv.submit(FakeCase(4, 90, 1000))  # 0: long-running, holds the budget
    srv.step()
    srv.submit(FakeCase(4, 60, 5))     # 1: never fits while 0 runs
    for _ in range(3):
        srv.step()                     # 1 passed over three times: now starved
    srv.submit(FakeCase(4, 10, 5))     # 2: would fit, but 1 is starving
    srv.step()
    assert {r.rid for r in srv.running.values()} == {0}
    assert [r.rid for r in srv.waiting] == [1, 2]


def test_oversized_request_rejected_at_submit():
    srv = Server(hbm_budget_bytes=100, batch=FakeBatch())
    with pytest.raises(ValueError):
        srv.submit(FakeCase(4, 101, 5))


def test_drain_finishes_everything():
    fb = FakeBatch()
    srv = Server(hbm_budget_bytes=0, chunk_steps=7, batch=fb)
    for i in range(5):
        srv.submit(FakeCase(4, 10, 3 + 11 * i))
    srv.drain()
    assert srv.stats.completed == 5 and not fb.m
    assert srv.stats.percentile(50) is not None and srv.stats.percentile(99) >= srv.stats.percentile(50)
