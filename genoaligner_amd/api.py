"""Python mirror of genoaligner's pairwise API (reference:
include/genoaligner/api.hpp:61-313, :485-487), over the C ABI of
libgenoaligner_amd.so.

Same names, same fields and defaults, same result rules and the same error
behaviour as the reference's C++ functions: a batch with any invalid request
comes back with ``status == Status.invalid_argument`` and default results, a
device failure with ``Status.device_error``; unresolved WFA pairs are score -1,
not errors.  Sequences may be ``str`` (latin-1) or bytes-like.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

from . import _native as N

SW_MAX_TRACE_CELLS = N.GA_SW_MAX_TRACE_CELLS


class Status(enum.Enum):
    ok = 0
    invalid_argument = 1
    device_error = 2


def _bytes(s) -> Optional[bytes]:
    if s is None:
        return None
    if isinstance(s, str):
        return s.encode("latin-1")
    return bytes(s)


@dataclass
class AlignRequest:
    pattern: Optional[object] = None
    pattern_len: int = 0
    text: Optional[object] = None
    text_len: int = 0
    smax: int = 64
    with_cigar: bool = True


@dataclass
class AlignResult:
    score: int = -1
    resolved: bool = False
    cigar: str = ""
    rescore_ok: bool = False
    wellformed_ok: bool = False


@dataclass
class BatchResult:
    results: List[AlignResult] = field(default_factory=list)
    resolved_count: int = 0
    unresolved_count: int = 0
    status: Status = Status.ok
    error: Optional[str] = None

    def ok(self) -> bool:
        return self.status == Status.ok


@dataclass
class SWScoring:
    match: int = 2
    mismatch: int = -3
    gap_open: int = 5
    gap_extend: int = 2


@dataclass
class SWRequest:
    text: Optional[object] = None
    text_len: int = 0
    pattern: Optional[object] = None
    pattern_len: int = 0
    scoring: SWScoring = field(default_factory=SWScoring)
    with_cigar: bool = True


@dataclass
class SWAlignResult:
    score: int = -1
    resolved: bool = False
    cigar: str = ""
    start_i: int = -1
    start_j: int = -1
    end_i: int = -1
    end_j: int = -1
    rescore_ok: bool = False
    wellformed_ok: bool = False
    too_large: bool = False


@dataclass
class SWBatchResult:
    results: List[SWAlignResult] = field(default_factory=list)
    resolved_count: int = 0
    unresolved_count: int = 0
    status: Status = Status.ok
    error: Optional[str] = None

    def ok(self) -> bool:
        return self.status == Status.ok


def _buf(b: Optional[bytes], keep: list) -> Optional[int]:
    """Address of a stable copy of ``b`` (None stays a null pointer)."""
    if b is None:
        return None
    cb = C.create_string_buffer(b, len(b) + 1)
    keep.append(cb)
    return C.addressof(cb)


def _wfa(reqs: Sequence[AlignRequest], ext: bool) -> BatchResult:
    n = len(reqs)
    out = BatchResult(results=[AlignResult() for _ in range(n)])
    if n == 0:
        return out
    keep: list = []
    arr = (N.AlignReq * n)()
    for i, r in enumerate(reqs):
        arr[i] = N.AlignReq(_buf(_bytes(r.pattern), keep), int(r.pattern_len),
                            _buf(_bytes(r.text), keep), int(r.text_len), int(r.smax),
                            1 if r.with_cigar else 0)
    res = (N.AlignRes * n)()
    L = N.lib()
    need = L.ga_cigar_arena_bytes_wfa(arr, n)
    arena = C.create_string_buffer(max(need, 1))
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    fn = L.ga_align_batch_ext if ext else L.ga_align_batch
    rc = fn(arr, n, res, arena, need, C.byref(rc_ok), C.byref(rc_un), C.byref(err))
    out.status = Status(rc)
    out.error = err.value.decode() if err.value else None
    if rc != N.GA_OK:
        return out
    raw = arena.raw
    for i in range(n):
        r, o = res[i], out.results[i]
        o.score, o.resolved = r.score, bool(r.resolved)
        if r.cigar_len > 0:
            o.cigar = raw[r.cigar_off: r.cigar_off + r.cigar_len].decode("latin-1")
        o.rescore_ok, o.wellformed_ok = bool(r.rescore_ok), bool(r.wellformed_ok)
    out.resolved_count, out.unresolved_count = rc_ok.value, rc_un.value
    return out


def align_batch(reqs: Sequence[AlignRequest]) -> BatchResult:
    """genoaligner::align_batch (api.hpp:164): one smax, the minimum over the
    batch, each in [0, 511]."""
    return _wfa(reqs, False)


def align_batch_ext(reqs: Sequence[AlignRequest]) -> BatchResult:
    """Extension: per-request smax in [0, 4000] and per-request with_cigar."""
    return _wfa(reqs, True)


def align(req: AlignRequest) -> AlignResult:
    b = align_batch([req])
    return b.results[0] if b.results else AlignResult()


def _scoring_c(s) -> N.SWScoringC:
    if isinstance(s, SWScoring):
        return N.SWScoringC(s.match, s.mismatch, s.gap_open, s.gap_extend)
    return N.SWScoringC(*[int(x) for x in s])


def _sw_call(reqs: Sequence[SWRequest], matrix=None) -> SWBatchResult:
    n = len(reqs)
    out = SWBatchResult(results=[SWAlignResult() for _ in range(n)])
    if n == 0:
        return out
    keep: list = []
    arr = (N.SWReq * n)()
    for i, r in enumerate(reqs):
        arr[i] = N.SWReq(_buf(_bytes(r.text), keep), int(r.text_len),
                         _buf(_bytes(r.pattern), keep), int(r.pattern_len),
                         _scoring_c(r.scoring), 1 if r.with_cigar else 0)
    res = (N.SWRes * n)()
    L = N.lib()
    need = L.ga_cigar_arena_bytes_sw(arr, n)
    arena = C.create_string_buffer(max(need, 1))
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    if matrix is None:
        rc = L.ga_align_sw_batch(arr, n, res, arena, need, C.byref(rc_ok), C.byref(rc_un), C.byref(err))
    else:
        rc = L.ga_align_sw_matrix_batch(arr, n, C.byref(matrix), res, arena, need, C.byref(rc_ok),
                                        C.byref(rc_un), C.byref(err))
    out.status = Status(rc)
    out.error = err.value.decode() if err.value else None
    if rc != N.GA_OK:
        return out
    raw = arena.raw
    for i in range(n):
        r, o = res[i], out.results[i]
        o.score, o.resolved = r.score, bool(r.resolved)
        if r.cigar_len > 0:
            o.cigar = raw[r.cigar_off: r.cigar_off + r.cigar_len].decode("latin-1")
        o.start_i, o.start_j, o.end_i, o.end_j = r.start_i, r.start_j, r.end_i, r.end_j
        o.rescore_ok, o.wellformed_ok = bool(r.rescore_ok), bool(r.wellformed_ok)
        o.too_large = bool(r.too_large)
    out.resolved_count, out.unresolved_count = rc_ok.value, rc_un.value
    return out


def align_sw_batch(reqs: Sequence[SWRequest]) -> SWBatchResult:
    """genoaligner::align_sw_batch (api.hpp:312)."""
    return _sw_call(reqs)


def blosum62(gap_open: int = 12, gap_extend: int = 1) -> "N.SwMatrix":
    """BLOSUM62 scheme (ARNDCQEGHILKMFPSTWYV, unknown residues -1).  A gap of
    length L costs gap_open + (L - 1) * gap_extend."""
    m = N.SwMatrix()
    N.lib().ga_sw_matrix_blosum62(C.byref(m), int(gap_open), int(gap_extend))
    return m


def align_sw_matrix_batch(reqs: Sequence[SWRequest], matrix) -> SWBatchResult:
    """Smith-Waterman under a substitution matrix (an extension: the same
    rules as align_sw_batch, with the requests' scoring replaced by `matrix`)."""
    return _sw_call(reqs, matrix)


def align_sw(req: SWRequest) -> SWAlignResult:
    b = align_sw_batch([req])
    return b.results[0] if b.results else SWAlignResult()


def backend_name() -> str:
    return N.lib().ga_backend_name().decode()


def device_name() -> str:
    return N.lib().ga_device_name().decode()


def device_available() -> bool:
    return bool(N.lib().ga_device_available())
