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

import numpy as np

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


def _np_dtype(cstruct) -> np.dtype:
    """numpy dtype with the field offsets and size of a ctypes struct (nested
    structs of int32 become int32 sub-arrays)."""
    names, formats, offsets = [], [], []
    for name, ctype in cstruct._fields_:
        off = getattr(cstruct, name).offset
        if ctype is C.c_void_p:
            fmt = np.uint64
        elif ctype is C.c_int32:
            fmt = np.int32
        elif ctype is C.c_int64:
            fmt = np.int64
        elif isinstance(ctype, type) and issubclass(ctype, C.Structure):
            fmt = (np.int32, C.sizeof(ctype) // 4)
        else:
            raise TypeError(f"no numpy layout for field {name}")
        names.append(name)
        formats.append(fmt)
        offsets.append(off)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": C.sizeof(cstruct)})


_ALIGN_REQ = _np_dtype(N.AlignReq)
_ALIGN_RES = _np_dtype(N.AlignRes)
_SW_REQ = _np_dtype(N.SWReq)
_SW_RES = _np_dtype(N.SWRes)


def _seq_table(seqs, lens, keep: list) -> np.ndarray:
    """Addresses (uint64, 0 for None) of the sequences, copied back to back into
    one buffer, each followed by zero bytes up to max(len, declared len) + 1."""
    bs = [_bytes(x) for x in seqs]
    parts, addr, off = [], np.zeros(len(bs), dtype=np.uint64), 0
    offs = []
    for b, ln in zip(bs, lens):
        if b is None:
            offs.append(-1)
            continue
        pad = max(int(ln) - len(b), 0) + 1
        parts.append(b)
        parts.append(bytes(pad))
        offs.append(off)
        off += len(b) + pad
    blob = b"".join(parts)
    keep.append(blob)
    base = C.cast(C.c_char_p(blob), C.c_void_p).value or 0
    o = np.asarray(offs, dtype=np.int64)
    addr[o >= 0] = np.uint64(base) + o[o >= 0].astype(np.uint64)
    return addr


def _ptr(a: np.ndarray, cstruct):
    return a.ctypes.data_as(C.POINTER(cstruct))


def _pack_rows(reqs, wfa: bool):
    """C request rows of `reqs` (ga_align_req when wfa, else ga_sw_req) and
    the buffers their sequence pointers address: (rows, keep)."""
    keep: list = []
    n = len(reqs)
    if wfa:
        arr = np.zeros(n, dtype=_ALIGN_REQ)
        plen = [int(r.pattern_len) for r in reqs]
        tlen = [int(r.text_len) for r in reqs]
        arr["pattern"] = _seq_table([r.pattern for r in reqs], plen, keep)
        arr["pattern_len"] = plen
        arr["text"] = _seq_table([r.text for r in reqs], tlen, keep)
        arr["text_len"] = tlen
        arr["smax"] = [int(r.smax) for r in reqs]
    else:
        arr = np.zeros(n, dtype=_SW_REQ)
        tlen = [int(r.text_len) for r in reqs]
        plen = [int(r.pattern_len) for r in reqs]
        arr["text"] = _seq_table([r.text for r in reqs], tlen, keep)
        arr["text_len"] = tlen
        arr["pattern"] = _seq_table([r.pattern for r in reqs], plen, keep)
        arr["pattern_len"] = plen
        arr["scoring"] = [(s.match, s.mismatch, s.gap_open, s.gap_extend) if isinstance(s, SWScoring)
                          else tuple(int(x) for x in s) for s in (r.scoring for r in reqs)]
    arr["with_cigar"] = [1 if r.with_cigar else 0 for r in reqs]
    return arr, keep


def _call_wfa_rows(arr: np.ndarray, ext: bool):
    """ga_align_batch(_ext) over packed rows: (result rows, CIGAR arena as a
    uint8 array, error or None)."""
    n = len(arr)
    res = np.zeros(n, dtype=_ALIGN_RES)
    L = N.lib()
    pa = _ptr(arr, N.AlignReq)
    need = L.ga_cigar_arena_bytes_wfa(pa, n)
    arena = np.empty(max(need, 1), dtype=np.uint8)
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    fn = L.ga_align_batch_ext if ext else L.ga_align_batch
    rc = fn(pa, n, _ptr(res, N.AlignRes), arena.ctypes.data_as(C.c_char_p), need, C.byref(rc_ok),
            C.byref(rc_un), C.byref(err))
    if rc != N.GA_OK:
        return res, arena, (err.value.decode() if err.value else f"status {rc}")
    return res, arena, None


def _call_sw_rows(arr: np.ndarray, matrix=None):
    """ga_align_sw_batch (or the matrix entry point) over packed rows."""
    n = len(arr)
    res = np.zeros(n, dtype=_SW_RES)
    L = N.lib()
    pa = _ptr(arr, N.SWReq)
    pr = _ptr(res, N.SWRes)
    need = L.ga_cigar_arena_bytes_sw(pa, n)
    arena = np.empty(max(need, 1), dtype=np.uint8)
    ap = arena.ctypes.data_as(C.c_char_p)
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    if matrix is None:
        rc = L.ga_align_sw_batch(pa, n, pr, ap, need, C.byref(rc_ok), C.byref(rc_un), C.byref(err))
    else:
        rc = L.ga_align_sw_matrix_batch(pa, n, C.byref(matrix), pr, ap, need, C.byref(rc_ok),
                                        C.byref(rc_un), C.byref(err))
    if rc != N.GA_OK:
        return res, arena, (err.value.decode() if err.value else f"status {rc}")
    return res, arena, None


def _align_results(res: np.ndarray, arena: np.ndarray) -> List[AlignResult]:
    raw = arena
    sc = res["score"].tolist()
    rv = res["resolved"].tolist()
    co = res["cigar_off"].tolist()
    cl = res["cigar_len"].tolist()
    rs = res["rescore_ok"].tolist()
    wf = res["wellformed_ok"].tolist()
    return [AlignResult(sc[i], bool(rv[i]),
                        raw[co[i]: co[i] + cl[i]].tobytes().decode("latin-1") if cl[i] > 0 else "",
                        bool(rs[i]), bool(wf[i])) for i in range(len(res))]


def _sw_results(res: np.ndarray, arena: np.ndarray) -> List["SWAlignResult"]:
    f = {k: res[k].tolist() for k in ("score", "resolved", "cigar_off", "cigar_len", "start_i", "start_j",
                                       "end_i", "end_j", "rescore_ok", "wellformed_ok", "too_large")}
    return [SWAlignResult(
        f["score"][i], bool(f["resolved"][i]),
        arena[f["cigar_off"][i]: f["cigar_off"][i] + f["cigar_len"][i]].tobytes().decode("latin-1")
        if f["cigar_len"][i] > 0 else "",
        f["start_i"][i], f["start_j"][i], f["end_i"][i], f["end_j"][i], bool(f["rescore_ok"][i]),
        bool(f["wellformed_ok"][i]), bool(f["too_large"][i])) for i in range(len(res))]


def _wfa(reqs: Sequence[AlignRequest], ext: bool) -> BatchResult:
    n = len(reqs)
    out = BatchResult(results=[])
    if n == 0:
        return out
    keep: list = []
    plen = [int(r.pattern_len) for r in reqs]
    tlen = [int(r.text_len) for r in reqs]
    arr = np.zeros(n, dtype=_ALIGN_REQ)
    arr["pattern"] = _seq_table([r.pattern for r in reqs], plen, keep)
    arr["pattern_len"] = plen
    arr["text"] = _seq_table([r.text for r in reqs], tlen, keep)
    arr["text_len"] = tlen
    arr["smax"] = [int(r.smax) for r in reqs]
    arr["with_cigar"] = [1 if r.with_cigar else 0 for r in reqs]
    res = np.zeros(n, dtype=_ALIGN_RES)
    L = N.lib()
    pa = _ptr(arr, N.AlignReq)
    need = L.ga_cigar_arena_bytes_wfa(pa, n)
    arena = C.create_string_buffer(max(need, 1))
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    fn = L.ga_align_batch_ext if ext else L.ga_align_batch
    rc = fn(pa, n, _ptr(res, N.AlignRes), arena, need, C.byref(rc_ok), C.byref(rc_un), C.byref(err))
    out.status = Status(rc)
    out.error = err.value.decode() if err.value else None
    if rc != N.GA_OK:
        out.results = [AlignResult() for _ in range(n)]
        return out
    raw = arena.raw
    sc = res["score"].tolist()
    rv = res["resolved"].tolist()
    co = res["cigar_off"].tolist()
    cl = res["cigar_len"].tolist()
    rs = res["rescore_ok"].tolist()
    wf = res["wellformed_ok"].tolist()
    out.results = [AlignResult(sc[i], bool(rv[i]),
                               raw[co[i]: co[i] + cl[i]].decode("latin-1") if cl[i] > 0 else "",
                               bool(rs[i]), bool(wf[i])) for i in range(n)]
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
    out = SWBatchResult(results=[])
    if n == 0:
        return out
    keep: list = []
    tlen = [int(r.text_len) for r in reqs]
    plen = [int(r.pattern_len) for r in reqs]
    arr = np.zeros(n, dtype=_SW_REQ)
    arr["text"] = _seq_table([r.text for r in reqs], tlen, keep)
    arr["text_len"] = tlen
    arr["pattern"] = _seq_table([r.pattern for r in reqs], plen, keep)
    arr["pattern_len"] = plen
    sc_of = [r.scoring for r in reqs]
    arr["scoring"] = [(s.match, s.mismatch, s.gap_open, s.gap_extend) if isinstance(s, SWScoring)
                      else tuple(int(x) for x in s) for s in sc_of]
    arr["with_cigar"] = [1 if r.with_cigar else 0 for r in reqs]
    res = np.zeros(n, dtype=_SW_RES)
    L = N.lib()
    pa = _ptr(arr, N.SWReq)
    pr = _ptr(res, N.SWRes)
    need = L.ga_cigar_arena_bytes_sw(pa, n)
    arena = C.create_string_buffer(max(need, 1))
    rc_ok, rc_un, err = C.c_int32(), C.c_int32(), C.c_char_p()
    if matrix is None:
        rc = L.ga_align_sw_batch(pa, n, pr, arena, need, C.byref(rc_ok), C.byref(rc_un), C.byref(err))
    else:
        rc = L.ga_align_sw_matrix_batch(pa, n, C.byref(matrix), pr, arena, need, C.byref(rc_ok),
                                        C.byref(rc_un), C.byref(err))
    out.status = Status(rc)
    out.error = err.value.decode() if err.value else None
    if rc != N.GA_OK:
        out.results = [SWAlignResult() for _ in range(n)]
        return out
    raw = arena.raw
    f = {k: res[k].tolist() for k in ("score", "resolved", "cigar_off", "cigar_len", "start_i", "start_j",
                                       "end_i", "end_j", "rescore_ok", "wellformed_ok", "too_large")}
    out.results = [SWAlignResult(
        f["score"][i], bool(f["resolved"][i]),
        raw[f["cigar_off"][i]: f["cigar_off"][i] + f["cigar_len"][i]].decode("latin-1") if f["cigar_len"][i] > 0
        else "",
        f["start_i"][i], f["start_j"][i], f["end_i"][i], f["end_j"][i], bool(f["rescore_ok"][i]),
        bool(f["wellformed_ok"][i]), bool(f["too_large"][i])) for i in range(n)]
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


def set_router(params: Optional[np.ndarray], q: float = 0.99) -> None:
    """Load a trained router network (GA_MLP_PARAMS float32 values; see
    genoaligner_amd.router) into the library, or unload it with None.  While
    one is loaded, align_batch and align_batch_ext run their WFA batches
    through the routed path (same results pair for pair)."""
    err = C.c_char_p()
    if params is None:
        rc = N.lib().ga_router_set(None, float(q), C.byref(err))
    else:
        arr = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if arr.size != N.GA_MLP_PARAMS:
            raise ValueError(f"router parameters: {arr.size} values, expected {N.GA_MLP_PARAMS}")
        rc = N.lib().ga_router_set(arr.ctypes.data, float(q), C.byref(err))
    if rc != N.GA_OK:
        raise RuntimeError((err.value or b"ga_router_set failed").decode())


def router_enabled() -> bool:
    return bool(N.lib().ga_router_enabled())


def align_sw(req: SWRequest) -> SWAlignResult:
    b = align_sw_batch([req])
    return b.results[0] if b.results else SWAlignResult()


def backend_name() -> str:
    return N.lib().ga_backend_name().decode()


def device_name() -> str:
    return N.lib().ga_device_name().decode()


def device_available() -> bool:
    return bool(N.lib().ga_device_available())
