"""Request loop with continuous batching over the library's batch entry points.

Requests (each 1..512 pairs; WFA DNA or substitution-matrix SW) arrive at any
time.  A fixed number of slots run batches concurrently, each on its own HIP
stream inside the library; whenever a slot frees up it takes, in arrival
order, every waiting request that is compatible with the oldest one and fits
the free memory budget, and runs them as one batch.  So requests that arrive
while others are on the GPU are admitted as soon as a slot and memory are free,
and their work overlaps the batches already running.

Memory budget: a batch reserves the device bytes of its sequences, CIGAR slots
and per-pair buffers (estimated from the lengths) plus the workspace cap its
slot runs under (ga_set_thread_workspace_budget); a request is admitted only
while the reserved total stays within the budget.

Results never depend on the mix: WFA batches use the per-pair entry point
(ga_align_batch_ext: each pair keeps its own smax and CIGAR flag), and SW
batches only join requests with the same CIGAR flag (the reference's SW rules
make a score-only pair's reported end cell depend on whether its batch traces).

Marshalling is per request, not per batch: at submission a request's pairs are
packed once (its sequences into one buffer, its C request rows pointing into
it); a slot joins the batch's rows and calls the C entry point directly, and
each request gets its slice of the result rows and the batch's CIGAR arena.
Result objects are built only when `Request.results` is read.
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import api
from . import _native as N


@dataclass
class Request:
    kind: str                       # "wfa" or "sw"
    items: list                     # AlignRequest (wfa) or SWRequest (sw)
    with_cigar: bool = True         # sw: the CIGAR flag shared by all items
    arrival: float = 0.0            # perf_counter() at submission
    admitted: float = 0.0
    done_at: float = 0.0
    error: Optional[str] = None
    rows: Optional[np.ndarray] = None    # packed C result rows (ga_align_res / ga_sw_res) of the items
    arena: Optional[np.ndarray] = None   # the CIGAR bytes their cigar_off / cigar_len address
    _results: Optional[list] = field(default=None, repr=False)
    _pack: Optional[tuple] = field(default=None, repr=False)
    _done: threading.Event = field(default_factory=threading.Event, repr=False)

    @property
    def results(self) -> Optional[list]:
        """Per-item result objects (built from the packed rows on first use)."""
        if self._results is None and self.rows is not None:
            self._results = (api._align_results if self.kind == "wfa" else api._sw_results)(
                self.rows, self.arena)
        return self._results

    @results.setter
    def results(self, value):
        self._results = value

    @property
    def latency_ms(self) -> float:
        return (self.done_at - self.arrival) * 1e3

    def wait(self, timeout: Optional[float] = None) -> bool:
        return self._done.wait(timeout)

    def key(self):
        return ("wfa",) if self.kind == "wfa" else ("sw", bool(self.with_cigar))

    def pairs(self) -> int:
        return len(self.items)

    def device_bytes(self) -> int:
        """Sequences, CIGAR slots and per-pair buffers on the device."""
        tot = 0
        for r in self.items:
            a = max(int(r.pattern_len), 0)
            b = max(int(r.text_len), 0)
            tot += 3 * (a + b) + 96
        return tot


def join_rows(parts: Sequence[np.ndarray]) -> np.ndarray:
    """The C request rows of several requests in one array of the same dtype
    (np.concatenate would drop the explicit field offsets and padding that
    match the C struct)."""
    if len(parts) == 1:
        return parts[0]
    rows = np.zeros(sum(len(p) for p in parts), dtype=parts[0].dtype)
    at = 0
    for p in parts:
        rows[at: at + len(p)] = p
        at += len(p)
    return rows


class LibraryBackend:
    """Runs one batch through the library (blocking; the library enqueues on a
    stream of its own and synchronises it before returning).  route=True loads
    the shipped router network into the library (genoaligner_amd.router.load),
    so WFA batches take the routed path; results are unchanged."""

    def __init__(self, matrix=None, route: bool = False, q: float = 0.99):
        self.matrix = matrix if matrix is not None else api.blosum62(11, 1)
        if route:
            from . import router
            api.set_router(router.load(), q)

    def set_workspace(self, nbytes: int):
        from . import _native as N
        N.lib().ga_set_thread_workspace_budget(int(nbytes))

    def release(self):
        """Free the idle workspace blocks the library keeps for reuse."""
        from . import _native as N
        N.lib().ga_trim_workspace()

    def run(self, key, items):
        if key[0] == "wfa":
            b = api.align_batch_ext(items)
        else:
            b = api.align_sw_matrix_batch(items, self.matrix)
        if not b.ok():
            raise RuntimeError(b.error or "batch failed")
        return b.results

    def pack(self, kind, items):
        """The request's C rows and the buffers they point into (built once,
        at submission)."""
        return api._pack_rows(items, kind == "wfa")

    def run_packed(self, key, packs):
        """One batch from packed requests: (result rows, CIGAR arena)."""
        rows = join_rows([p[0] for p in packs])
        if key[0] == "wfa":
            res, arena, err = api._call_wfa_rows(rows, True)
        else:
            res, arena, err = api._call_sw_rows(rows, self.matrix)
        if err is not None:
            raise RuntimeError(err)
        return res, arena


class Server:
    def __init__(self, budget_bytes: int = 64 << 30, slots: int = 4, max_batch_pairs: int = 16384,
                 backend=None, clock: Callable[[], float] = time.perf_counter, start: bool = True):
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.workspace = self.budget // (2 * self.slots)     # per running batch
        self.max_batch_pairs = int(max_batch_pairs)
        self.backend = backend if backend is not None else LibraryBackend()
        self.clock = clock
        self._cv = threading.Condition()
        self._pending: collections.deque = collections.deque()
        self._reserved = 0
        self._closed = False
        self.peak_reserved = 0
        self.batches: List[dict] = []                          # one record per batch run
        self._threads = [threading.Thread(target=self._slot, name=f"ga-slot-{i}", daemon=True)
                         for i in range(self.slots)]
        self._started = False
        if start:
            self.start()

    def start(self):
        """Start the slots (a server created with start=False only queues)."""
        if not self._started:
            self._started = True
            for t in self._threads:
                t.start()

    # ------------------------------------------------------------ client side
    def submit(self, kind: str, items: Sequence, with_cigar: bool = True) -> Request:
        req = Request(kind=kind, items=list(items), with_cigar=with_cigar, arrival=self.clock())
        pack = getattr(self.backend, "pack", None)
        if pack is not None and req.items:
            req._pack = pack(kind, req.items)
        refuse = None
        if kind == "sw" and any(bool(it.with_cigar) != bool(with_cigar) for it in req.items):
            # SW batches are keyed by the flag, and a score-only pair's reported
            # end cell depends on whether its batch traces: a request whose
            # items disagree with its flag could land in a mixed batch
            refuse = "SW request items must all have with_cigar equal to the request's flag"
        elif req.device_bytes() + self.workspace > self.budget:
            refuse = "request larger than the memory budget"
        if refuse:
            req.error = refuse
            req.done_at = self.clock()
            req._done.set()
            return req
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def close(self, wait: bool = True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait and self._started:
            for t in self._threads:
                t.join()
            release = getattr(self.backend, "release", None)
            if release is not None:
                release()

    # ------------------------------------------------------------ slots
    def _take(self):
        """Under the lock: the oldest request and every later compatible one
        that fits the free budget and the pair cap; None if the oldest does
        not fit yet."""
        head = self._pending[0]
        key = head.key()
        need = head.device_bytes() + self.workspace
        if self._reserved + need > self.budget:
            return None
        batch, pairs, used = [head], head.pairs(), need
        for r in list(self._pending)[1:]:
            if r.key() != key:
                continue
            b = r.device_bytes()
            if pairs + r.pairs() > self.max_batch_pairs or self._reserved + used + b > self.budget:
                continue
            batch.append(r)
            pairs += r.pairs()
            used += b
        for r in batch:
            self._pending.remove(r)
        self._reserved += used
        self.peak_reserved = max(self.peak_reserved, self._reserved)
        return key, batch, used

    def _slot(self):
        if hasattr(self.backend, "set_workspace"):
            self.backend.set_workspace(self.workspace)
        while True:
            with self._cv:
                took = None
                while True:
                    if self._pending:
                        took = self._take()
                        if took:
                            break
                    elif self._closed:
                        return
                    self._cv.wait(0.05)
            key, batch, used = took
            t0 = self.clock()
            for r in batch:
                r.admitted = t0
            items = [it for r in batch for it in r.items]
            err = None
            packed = all(r._pack is not None for r in batch)
            try:
                if packed:
                    res, arena = self.backend.run_packed(key, [r._pack for r in batch])
                else:
                    res, arena = self.backend.run(key, items), None
            except Exception as e:                      # reported per request
                res, arena, err = None, None, str(e)
            t1 = self.clock()
            at = 0
            for r in batch:
                if res is not None:
                    if packed:
                        r.rows, r.arena = res[at: at + r.pairs()], arena
                    else:
                        r.results = res[at: at + r.pairs()]
                r.error = err
                at += r.pairs()
                r.done_at = t1
            with self._cv:
                self._reserved -= used
                self.batches.append({"key": key, "requests": len(batch), "pairs": len(items),
                                     "start": t0, "end": t1, "bytes": used})
                self._cv.notify_all()
            for r in batch:
                r._done.set()


def poisson_load(server: Server, make_request: Callable[[int], tuple], rate_per_s: float, count: int,
                 seed: int = 1) -> List[Request]:
    """Submit `count` requests with exponential inter-arrival times (rate
    `rate_per_s`); make_request(k) -> (kind, items, with_cigar)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count)
    out = []
    t = time.perf_counter()
    for k in range(count):
        t += gaps[k]
        now = time.perf_counter()
        if t > now:
            time.sleep(t - now)
        kind, items, cig = make_request(k)
        out.append(server.submit(kind, items, with_cigar=cig))
    return out
