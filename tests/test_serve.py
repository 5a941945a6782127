"""Scheduling logic of the request loop, on the CPU with a stand-in backend:
continuous batching, memory-budget admission, compatibility keys, per-request
result slicing and failure reporting."""
import threading
import time

import pytest

from genoaligner_amd import api
from genoaligner_amd.serve import Server


class FakeBackend:
    """Returns, per item, (batch id, item tag); sleeps to keep batches in flight."""

    def __init__(self, delay=0.05, fail_on=None):
        self.delay = delay
        self.fail_on = fail_on
        self.lock = threading.Lock()
        self.active = 0
        self.peak = 0
        self.calls = []

    def run(self, key, items):
        with self.lock:
            self.active += 1
            self.peak = max(self.peak, self.active)
            bid = len(self.calls)
            self.calls.append((key, [it.tag for it in items]))
        try:
            time.sleep(self.delay)
            if self.fail_on and any(it.tag == self.fail_on for it in items):
                raise RuntimeError("planted failure")
            return [(bid, it.tag) for it in items]
        finally:
            with self.lock:
                self.active -= 1


def _wfa(tag, length=100):
    r = api.AlignRequest(pattern="A" * length, pattern_len=length, text="A" * length, text_len=length,
                         smax=10, with_cigar=True)
    r.tag = tag
    return r


def _sw(tag, length=100, cigar=True):
    r = api.SWRequest(text="A" * length, text_len=length, pattern="A" * length, pattern_len=length,
                      with_cigar=cigar)
    r.tag = tag
    return r


def test_results_are_sliced_per_request_in_order():
    fb = FakeBackend(delay=0.01)
    s = Server(budget_bytes=1 << 30, slots=2, backend=fb)
    reqs = [s.submit("wfa", [_wfa(f"r{k}i{i}") for i in range(k + 1)]) for k in range(6)]
    for r in reqs:
        assert r.wait(5)
    s.close()
    for k, r in enumerate(reqs):
        assert r.error is None
        assert [tag for _, tag in r.results] == [f"r{k}i{i}" for i in range(k + 1)]


def test_continuous_batching_admits_arrivals_while_busy():
    """Requests that arrive while every slot is busy are batched together as
    soon as a slot frees up."""
    fb = FakeBackend(delay=0.2)
    s = Server(budget_bytes=1 << 30, slots=1, backend=fb)
    first = s.submit("wfa", [_wfa("a")])
    time.sleep(0.05)                                  # the first batch is running
    later = [s.submit("wfa", [_wfa(f"b{k}")]) for k in range(5)]
    for r in [first] + later:
        assert r.wait(5)
    s.close()
    assert fb.calls[0][1] == ["a"]
    assert fb.calls[1][1] == [f"b{k}" for k in range(5)]   # one batch for all five
    assert len({r.results[0][0] for r in later}) == 1


def test_slots_overlap():
    fb = FakeBackend(delay=0.2)
    s = Server(budget_bytes=1 << 30, slots=3, max_batch_pairs=1, backend=fb)
    reqs = [s.submit("wfa", [_wfa(f"x{k}")]) for k in range(3)]
    for r in reqs:
        assert r.wait(5)
    s.close()
    assert fb.peak == 3


def test_memory_budget_is_never_exceeded():
    fb = FakeBackend(delay=0.05)
    one = _wfa("probe", 1000)
    from genoaligner_amd.serve import Request
    per = Request(kind="wfa", items=[one]).device_bytes()
    slots = 4
    budget = 8 * per * 2 * slots                      # workspace share = 8 requests' bytes
    s = Server(budget_bytes=budget, slots=slots, backend=fb)
    reqs = [s.submit("wfa", [_wfa(f"m{k}", 1000)]) for k in range(40)]
    for r in reqs:
        assert r.wait(10)
    s.close()
    assert s.peak_reserved <= budget
    assert all(b["bytes"] <= budget for b in s.batches)
    assert fb.peak <= slots


def test_incompatible_requests_are_not_mixed():
    fb = FakeBackend(delay=0.1)
    s = Server(budget_bytes=1 << 30, slots=1, backend=fb)
    blocker = s.submit("wfa", [_wfa("block")])
    time.sleep(0.03)
    a = s.submit("sw", [_sw("s1", cigar=True)], with_cigar=True)
    b = s.submit("sw", [_sw("s2", cigar=False)], with_cigar=False)
    c = s.submit("wfa", [_wfa("w1")])
    d = s.submit("sw", [_sw("s3", cigar=True)], with_cigar=True)
    for r in (blocker, a, b, c, d):
        assert r.wait(5)
    s.close()
    groups = [tuple(tags) for _, tags in fb.calls]
    assert ("s1", "s3") in groups and ("s2",) in groups and ("w1",) in groups


def test_failures_are_reported_per_request():
    fb = FakeBackend(delay=0.01, fail_on="bad")
    s = Server(budget_bytes=1 << 30, slots=1, max_batch_pairs=1, backend=fb)
    ok = s.submit("wfa", [_wfa("good")])
    bad = s.submit("wfa", [_wfa("bad")])
    for r in (ok, bad):
        assert r.wait(5)
    s.close()
    assert ok.error is None and ok.results
    assert bad.error == "planted failure" and bad.results is None


def test_oversized_request_is_refused():
    s = Server(budget_bytes=10_000, slots=1, backend=FakeBackend())
    r = s.submit("wfa", [_wfa("huge", 10_000)])
    assert r.wait(1) and r.error == "request larger than the memory budget"
    s.close()


def test_paused_server_batches_the_queue():
    fb = FakeBackend(delay=0.01)
    s = Server(budget_bytes=1 << 30, slots=2, backend=fb, start=False)
    reqs = [s.submit("wfa", [_wfa(f"q{k}")]) for k in range(8)]
    s.start()
    for r in reqs:
        assert r.wait(5)
    s.close()
    assert len(fb.calls) <= 2 and sum(len(t) for _, t in fb.calls) == 8


def test_sw_request_with_disagreeing_item_flags_is_refused():
    """SW batches are keyed by the request's CIGAR flag; an item whose own flag
    differs would make a mixed batch, so such a request is refused."""
    fb = FakeBackend(delay=0.01)
    s = Server(budget_bytes=1 << 30, slots=1, backend=fb)
    r = s.submit("sw", [_sw("a", cigar=True), _sw("b", cigar=False)], with_cigar=True)
    ok = s.submit("sw", [_sw("c", cigar=False)], with_cigar=False)
    assert r.wait(1) and r.error.startswith("SW request items must all have with_cigar")
    assert ok.wait(5) and ok.error is None
    s.close()
    assert [tags for _, tags in fb.calls] == [["c"]]


def test_joined_request_rows_keep_the_c_layout():
    """A batch's packed rows (several requests joined) keep the C structs'
    field offsets and size, and every row keeps its request's values."""
    import ctypes as C
    import numpy as np
    from genoaligner_amd import _native as N
    from genoaligner_amd.serve import join_rows
    for wfa, cs in ((True, N.AlignReq), (False, N.SWReq)):
        parts = []
        for k in range(3):
            items = [(_wfa if wfa else _sw)(f"r{k}i{i}", length=10 + 7 * k + i) for i in range(k + 2)]
            parts.append(api._pack_rows(items, wfa)[0])
        rows = join_rows(parts)
        assert rows.dtype == parts[0].dtype and rows.dtype.itemsize == C.sizeof(cs)
        for name, _ in cs._fields_:
            assert rows.dtype.fields[name][1] == getattr(cs, name).offset, name
        at = 0
        for p in parts:
            assert rows[at: at + len(p)].tobytes() == p.tobytes()
            at += len(p)
