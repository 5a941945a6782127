"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module.  It loads

* ``oracle/_build/libga_oracle.so`` -- the C restatement (oracle/ga_oracle.c) of
  the reference's per-pair algorithms, and
* ``oracle/_ref/libga_ref.so`` (optional) -- the reference's own kernel bodies
  and API compiled for the CPU through its HIP CPU shim (oracle/build_ref.sh),

and restates the batch rules of the reference API (src/api/api.cpp:234-497 for
WFA, :584-847 for Smith-Waterman) on top of the per-pair functions, so a test
can ask "what would the reference return for this whole batch".
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libga_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libga_ref.so")

# api.cpp:227 and api.hpp:256
K_MAX_SMAX = 511
SW_MAX_TRACE_CELLS = 1 << 26

_lib = None
_ref = None


def build(quiet: bool = True) -> None:
    """Compile the C restatement (and the reference shim library when the
    reference tree is present)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=out)
    subprocess.check_call([os.path.join(HERE, "build_ref.sh")], stdout=out, stderr=out)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        cp, i32p = C.c_char_p, C.POINTER(C.c_int)
        L.ga_oracle_wfa.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int, i32p, C.c_void_p, i32p]
        L.ga_oracle_wfa_score.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int]
        L.ga_oracle_levenshtein.argtypes = [cp, C.c_int, cp, C.c_int]
        L.ga_oracle_wfa_validate.argtypes = [cp, C.c_int, C.c_int, cp, C.c_int, cp, C.c_int, i32p, i32p]
        L.ga_oracle_sw_table_mm.argtypes = [C.c_int, C.c_int, C.c_void_p]
        sw_args = [cp, C.c_int, cp, C.c_int, C.c_void_p, C.c_int, C.c_int,
                   i32p, i32p, i32p, i32p, i32p, C.c_void_p, i32p]
        L.ga_oracle_sw_trace.argtypes = sw_args
        L.ga_oracle_sw_value_walk.argtypes = sw_args
        L.ga_oracle_sw_score_lane.argtypes = [cp, C.c_int, cp, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                              C.c_int, i32p, i32p, i32p]
        L.ga_oracle_sw_validate.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int, C.c_int,
                                            cp, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                            C.c_int, C.c_int, C.c_int, i32p, i32p]
        vp = C.c_void_p
        L.ga_oracle_wfa_batch.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int64,
                                          vp, vp, vp, vp, C.c_int]
        L.ga_oracle_sw_batch.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, vp, C.c_int, C.c_int,
                                         C.c_int, vp, vp, vp, vp, C.c_int]
        L.ga_oracle_sketch.argtypes = [cp, C.c_int, cp, C.c_int, vp, vp]
        L.ga_oracle_sketch.restype = None
        L.ga_oracle_mlp_forward.argtypes = [vp, vp, C.c_int64, vp, vp, vp]
        L.ga_oracle_mlp_forward.restype = None
        L.ga_oracle_mlp_backward.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, vp, vp]
        L.ga_oracle_mlp_backward.restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own code compiled for the CPU (None when not built)."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        R = C.CDLL(REF_SO)
        cp, i32p, vp = C.c_char_p, C.POINTER(C.c_int), C.c_void_p
        R.ga_ref_wfa_trace.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int, i32p, vp, i32p, i32p, i32p]
        R.ga_ref_wfa_score_flat.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int]
        R.ga_ref_edit_distance.argtypes = [cp, C.c_int, cp, C.c_int]
        R.ga_ref_sw_trace.argtypes = [cp, C.c_int, cp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      vp, vp, i32p]
        R.ga_ref_sw_score.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, vp]
        R.ga_ref_align_sw_batch.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int,
                                            vp, vp, vp, vp]
        R.ga_ref_align_batch_status.argtypes = [vp, vp, vp, vp, vp, C.c_int]
        R.ga_ref_fasta_parse.argtypes = [cp, C.c_longlong, C.POINTER(C.c_longlong), vp, C.c_int]
        R.ga_ref_fasta_parse.restype = C.c_longlong
        R.ga_ref_fasta_blob.argtypes = [vp]
        R.ga_ref_fasta_blob.restype = None
        R.ga_ref_fasta_file.argtypes = [cp, vp, C.c_int]
        R.ga_ref_fasta_file.restype = C.c_longlong
        _ref = R
    return _ref


def _b(s) -> bytes:
    return s if isinstance(s, (bytes, bytearray)) else s.encode("latin-1")


# ---------------------------------------------------------------------------
# per-pair WFA
# ---------------------------------------------------------------------------
def wfa(pattern, text, smax: int):
    """(score, cigar) as wfa_trace_kernel + decode_cigar produce them; cigar is
    '' when unresolved or when the walk emitted nothing."""
    p, t = _b(pattern), _b(text)
    sc, ln = C.c_int(), C.c_int()
    buf = C.create_string_buffer(len(p) + len(t) + 4)
    if lib().ga_oracle_wfa(p, len(p), t, len(t), smax, C.byref(sc), buf, C.byref(ln)) != 0:
        raise MemoryError("oracle wfa")
    return sc.value, buf.raw[: ln.value].decode("latin-1")


def wfa_score(pattern, text, smax: int) -> int:
    p, t = _b(pattern), _b(text)
    return lib().ga_oracle_wfa_score(p, len(p), t, len(t), smax)


def levenshtein(pattern, text) -> int:
    p, t = _b(pattern), _b(text)
    return lib().ga_oracle_levenshtein(p, len(p), t, len(t))


def wfa_validate(cigar: str, score: int, pattern, text):
    p, t, c = _b(pattern), _b(text), _b(cigar)
    r, w = C.c_int(), C.c_int()
    lib().ga_oracle_wfa_validate(c, len(c), score, p, len(p), t, len(t), C.byref(r), C.byref(w))
    return bool(r.value), bool(w.value)


# ---------------------------------------------------------------------------
# per-pair SW
# ---------------------------------------------------------------------------
def mm_table(match: int, mismatch: int) -> np.ndarray:
    tab = np.empty(256 * 256, dtype=np.int32)
    lib().ga_oracle_sw_table_mm(match, mismatch, tab.ctypes.data)
    return tab


def matrix_table(alphabet: str, score, unknown: int) -> np.ndarray:
    """256x256 table (text byte, pattern byte) of a substitution matrix over
    `alphabet`; any pair with a byte outside it scores `unknown`."""
    tab = np.full((256, 256), int(unknown), dtype=np.int32)
    idx = [ord(c) for c in alphabet]
    sc = np.asarray(score, dtype=np.int32)[: len(idx), : len(idx)]
    tab[np.ix_(idx, idx)] = sc
    return tab.reshape(-1)


def _table(scoring) -> np.ndarray:
    if isinstance(scoring, np.ndarray):
        return np.ascontiguousarray(scoring, dtype=np.int32).reshape(-1)
    return mm_table(scoring[0], scoring[1])


def _gaps(scoring, gap_open=None, gap_extend=None):
    if isinstance(scoring, np.ndarray):
        return gap_open, gap_extend
    return scoring[2], scoring[3]


def sw_trace(text, pattern, scoring, gap_open=None, gap_extend=None):
    """sw_trace_kernel + sw_trace_walk: (score, start_i, start_j, end_i, end_j, cigar).
    ``scoring`` is (match, mismatch, gap_open, gap_extend) or a 256x256 table
    (then pass gap_open/gap_extend)."""
    t, p = _b(text), _b(pattern)
    tab = _table(scoring)
    go, ge = _gaps(scoring, gap_open, gap_extend)
    o = [C.c_int() for _ in range(6)]
    buf = C.create_string_buffer(len(t) + len(p) + 4)
    if lib().ga_oracle_sw_trace(t, len(t), p, len(p), tab.ctypes.data, go, ge,
                                *[C.byref(x) for x in o[:5]], buf, C.byref(o[5])) != 0:
        raise MemoryError("oracle sw_trace")
    return (o[0].value, o[1].value, o[2].value, o[3].value, o[4].value,
            buf.raw[: o[5].value].decode("latin-1"))


def sw_value_walk(text, pattern, scoring, gap_open=None, gap_extend=None):
    t, p = _b(text), _b(pattern)
    tab = _table(scoring)
    go, ge = _gaps(scoring, gap_open, gap_extend)
    o = [C.c_int() for _ in range(6)]
    buf = C.create_string_buffer(len(t) + len(p) + 4)
    if lib().ga_oracle_sw_value_walk(t, len(t), p, len(p), tab.ctypes.data, go, ge,
                                     *[C.byref(x) for x in o[:5]], buf, C.byref(o[5])) != 0:
        raise MemoryError("oracle sw_value_walk")
    return (o[0].value, o[1].value, o[2].value, o[3].value, o[4].value,
            buf.raw[: o[5].value].decode("latin-1"))


def sw_score_lane(text, pattern, scoring, gap_open=None, gap_extend=None, wsize: int = 64):
    """sw_score_kernel's (score, end_i, end_j) at warp width ``wsize``."""
    t, p = _b(text), _b(pattern)
    tab = _table(scoring)
    go, ge = _gaps(scoring, gap_open, gap_extend)
    s, i, j = C.c_int(), C.c_int(), C.c_int()
    lib().ga_oracle_sw_score_lane(t, len(t), p, len(p), tab.ctypes.data, go, ge, wsize,
                                  C.byref(s), C.byref(i), C.byref(j))
    return s.value, i.value, j.value


def sw_validate(cigar, text, start_i, end_i, pattern, start_j, end_j, scoring, score,
                gap_open=None, gap_extend=None):
    t, p, c = _b(text), _b(pattern), _b(cigar)
    tab = _table(scoring)
    go, ge = _gaps(scoring, gap_open, gap_extend)
    r, w = C.c_int(), C.c_int()
    lib().ga_oracle_sw_validate(c, len(c), t, len(t), start_i, end_i, p, len(p), start_j, end_j,
                                tab.ctypes.data, go, ge, score, C.byref(r), C.byref(w))
    return bool(r.value), bool(w.value)


# ---------------------------------------------------------------------------
# batch rules of the reference API, restated over the per-pair functions
# ---------------------------------------------------------------------------
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
    status: str = "ok"           # ok | invalid_argument | device_error
    error: Optional[str] = None


def align_batch(reqs: Sequence) -> BatchResult:
    """src/api/api.cpp:234-497.  Each request has .pattern .pattern_len .text
    .text_len .smax .with_cigar (pattern/text may be None)."""
    out = BatchResult(results=[AlignResult() for _ in reqs])
    if not reqs:
        return out
    for r in reqs:                                   # api.cpp:246-266
        if (r.pattern_len > 0 and r.pattern is None) or (r.text_len > 0 and r.text is None):
            return BatchResult([AlignResult() for _ in reqs], status="invalid_argument",
                               error="null pointer with non-zero length")
        if r.pattern_len < 0 or r.text_len < 0:
            return BatchResult([AlignResult() for _ in reqs], status="invalid_argument",
                               error="negative length")
        if r.smax < 0 or r.smax > K_MAX_SMAX:
            return BatchResult([AlignResult() for _ in reqs], status="invalid_argument",
                               error="smax out of range [0, 511]")
    smax = min(r.smax for r in reqs)                 # api.cpp:271-272
    any_cigar = any(r.with_cigar for r in reqs)      # api.cpp:273-274
    for k, r in enumerate(reqs):
        p = _b(r.pattern or b"")[: r.pattern_len]
        t = _b(r.text or b"")[: r.text_len]
        res = out.results[k]
        if any_cigar:                                # api.cpp:456-469
            sc, cg = wfa(p, t, smax)
            res.score, res.resolved = sc, sc >= 0
            if res.resolved and len(cg) > 0:
                res.cigar = cg
                res.rescore_ok, res.wellformed_ok = wfa_validate(cg, sc, p, t)
        else:                                        # api.cpp:471-488
            sc = wfa_score(p, t, smax)
            res.score, res.resolved = sc, sc >= 0
    out.resolved_count = sum(1 for r in out.results if r.resolved)
    out.unresolved_count = len(out.results) - out.resolved_count
    return out


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
    status: str = "ok"
    error: Optional[str] = None


def sw_scheme_error(sc, with_cigar: bool) -> Optional[str]:
    """api.cpp:515-536"""
    match, mismatch, go, ge = sc
    if match <= 0:
        return "SW scoring: match must be > 0 (the score is 0 for every input otherwise)"
    if mismatch >= match:
        return "SW scoring: mismatch must be < match"
    if go <= 0 or ge <= 0:
        return "SW scoring: gap costs must be > 0"
    if with_cigar and ge > go:
        return "SW scoring: gap_extend > gap_open with with_cigar"
    return None


def _sc4(sc):
    """A scoring scheme as a 4-tuple (accepts tuples or objects with fields)."""
    if hasattr(sc, "match"):
        return (sc.match, sc.mismatch, sc.gap_open, sc.gap_extend)
    return tuple(sc)


def align_sw_batch(reqs: Sequence) -> SWBatchResult:
    """src/api/api.cpp:584-847 (device path semantics).  Each request has .text
    .text_len .pattern .pattern_len .scoring (4-tuple) .with_cigar."""
    n = len(reqs)
    out = SWBatchResult(results=[SWAlignResult() for _ in reqs])
    if n == 0:
        return out

    def invalid(msg):
        return SWBatchResult([SWAlignResult() for _ in reqs], status="invalid_argument", error=msg)

    s0 = _sc4(reqs[0].scoring)
    for r in reqs:                                   # api.cpp:600-616
        if (r.text_len > 0 and r.text is None) or (r.pattern_len > 0 and r.pattern is None):
            return invalid("null pointer with non-zero length")
        if r.text_len < 0 or r.pattern_len < 0:
            return invalid("negative length")
        e = sw_scheme_error(_sc4(r.scoring), r.with_cigar)
        if e:
            return invalid(e)
        if _sc4(r.scoring) != s0:
            return invalid("all requests in an SW batch must share one scoring scheme")
    any_cigar = any(r.with_cigar for r in reqs)
    max_cells = max((r.text_len + 1) * (r.pattern_len + 1) for r in reqs)
    dir_stride = min(max_cells, SW_MAX_TRACE_CELLS)  # api.cpp:644-645
    for k, r in enumerate(reqs):
        t = _b(r.text or b"")[: r.text_len]
        p = _b(r.pattern or b"")[: r.pattern_len]
        res = out.results[k]
        if any_cigar:                                # api.cpp:750-784, 542-580
            m, nn = len(t), len(p)
            if m > 0 and nn > 0 and (m + 1) * (nn + 1) > dir_stride:
                res.too_large = True
                continue
            sc, si, sj, ei, ej, cg = sw_trace(t, p, s0)
            res.score, res.resolved, res.end_i, res.end_j = sc, True, ei, ej
            if sc == 0:
                res.rescore_ok = res.wellformed_ok = True
                continue
            if not r.with_cigar:
                continue
            res.start_i, res.start_j = si, sj
            if len(cg) > 0:
                res.cigar = cg
                res.rescore_ok, res.wellformed_ok = sw_validate(cg, t, si, ei, p, sj, ej, s0, sc)
        else:                                        # api.cpp:785-838
            sc, ei, ej = sw_score_lane(t, p, s0, wsize=64)
            res.score, res.resolved, res.end_i, res.end_j = sc, True, ei, ej
            if sc == 0:
                res.rescore_ok = res.wellformed_ok = True
    out.resolved_count = sum(1 for r in out.results if r.resolved)
    out.unresolved_count = n - out.resolved_count
    return out


# ---------------------------------------------------------------------------
# reference (compiled through its CPU shim) -- available only where built
# ---------------------------------------------------------------------------
def ref_wfa(pattern, text, smax: int):
    R = ref()
    p, t = _b(pattern), _b(text)
    sc, ln, em, tr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    buf = C.create_string_buffer(len(p) + len(t) + 4)
    R.ga_ref_wfa_trace(p, len(p), t, len(t), smax, C.byref(sc), buf, C.byref(ln), C.byref(em),
                       C.byref(tr))
    return sc.value, buf.raw[: ln.value].decode("latin-1")


def ref_wfa_score_flat(pattern, text, smax: int) -> int:
    p, t = _b(pattern), _b(text)
    return ref().ga_ref_wfa_score_flat(p, len(p), t, len(t), smax)


def ref_sw_trace(text, pattern, scoring):
    t, p = _b(text), _b(pattern)
    o = (C.c_int * 5)()
    ln = C.c_int()
    buf = C.create_string_buffer(len(t) + len(p) + 4)
    ref().ga_ref_sw_trace(t, len(t), p, len(p), *scoring, o, buf, C.byref(ln))
    return (o[0], o[1], o[2], o[3], o[4], buf.raw[: ln.value].decode("latin-1"))


def ref_sw_score(pairs, scoring, wsize: int = 64):
    """sw_score_kernel under the shim's warp emulation; pairs = [(text, pattern)]."""
    texts = [_b(t) for t, _ in pairs]
    pats = [_b(p) for _, p in pairs]
    T = b"".join(texts) or b"\0"
    P = b"".join(pats) or b"\0"
    toff = np.cumsum([0] + [len(x) for x in texts[:-1]]).astype(np.int64)
    poff = np.cumsum([0] + [len(x) for x in pats[:-1]]).astype(np.int64)
    tl = np.array([len(x) for x in texts], dtype=np.int32)
    pl = np.array([len(x) for x in pats], dtype=np.int32)
    out = np.zeros(3 * len(pairs), dtype=np.int32)
    Tb = C.create_string_buffer(T, len(T))
    Pb = C.create_string_buffer(P, len(P))
    ref().ga_ref_sw_score(C.addressof(Tb), toff.ctypes.data, tl.ctypes.data, C.addressof(Pb),
                          poff.ctypes.data, pl.ctypes.data, len(pairs), *scoring, wsize,
                          out.ctypes.data)
    return [tuple(int(v) for v in out[3 * i: 3 * i + 3]) for i in range(len(pairs))]


# ------------------------------------------------------------------ router
# (no reference counterpart: restatements of this build's own definitions,
# pinned in tests/test_router_oracle.py)
SKETCH_DIM, MLP_HIDDEN, ROUTE_BUCKETS, MLP_PARAMS = 64, 128, 16, 26896


def sketch(pattern, text):
    """(features float32[64], lower bound) of one pair."""
    p, t = _b(pattern), _b(text)
    f = np.zeros(SKETCH_DIM, dtype=np.float32)
    lb = C.c_int32()
    lib().ga_oracle_sketch(p, len(p), t, len(t), f.ctypes.data, C.byref(lb))
    return f, lb.value


def mlp_forward(params: np.ndarray, x: np.ndarray):
    """(h1, h2, logits) for rows x (float32 [n, 64])."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    h1 = np.zeros((n, MLP_HIDDEN), np.float32)
    h2 = np.zeros((n, MLP_HIDDEN), np.float32)
    lg = np.zeros((n, ROUTE_BUCKETS), np.float32)
    lib().ga_oracle_mlp_forward(params.ctypes.data, x.ctypes.data, n, h1.ctypes.data, h2.ctypes.data,
                                lg.ctypes.data)
    return h1, h2, lg


def mlp_backward(params: np.ndarray, x: np.ndarray, label: np.ndarray):
    """(grads float32[26896], per-sample loss) of the mean cross-entropy."""
    h1, h2, lg = mlp_forward(params, x)
    label = np.ascontiguousarray(label, dtype=np.int32)
    n = x.shape[0]
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(n, np.float32)
    lib().ga_oracle_mlp_backward(np.ascontiguousarray(params, dtype=np.float32).ctypes.data,
                                 np.ascontiguousarray(x, dtype=np.float32).ctypes.data, h1.ctypes.data,
                                 h2.ctypes.data, lg.ctypes.data, label.ctypes.data, n, g.ctypes.data,
                                 loss.ctypes.data)
    return g, loss


# ------------------------------------------------------------------ FASTA
def fasta_parse(data: bytes):
    """(records [(id, description, sequence)], None) or (None, error) for an
    in-memory FASTA file: a restatement of io::read_fasta
    (include/genoaligner/io/fasta.hpp:59-157), a byte scan with the same states
    (between records / inside a record), line counting and messages."""
    data = _b(data)
    total, i, line = len(data), 0, 1
    nl, cr, sp, tab, gt = 10, 13, 32, 9, 62

    nxt = {b"\n": -1, b"\r": -1}      # next position of each line-break byte seen so far

    def eol(j):                       # first \n or \r at or after j (total if none)
        for ch in nxt:
            if nxt[ch] < j <= total:
                f = data.find(ch, j)
                nxt[ch] = total if f < 0 else f
        return min(nxt.values())

    def consume_eol(j, ln):           # fasta.hpp:86-98
        e = eol(j)
        if e == total:
            return total, ln
        if data[e] == cr and e + 1 < total and data[e + 1] == nl:
            return e + 2, ln + 1
        return e + 1, ln + 1

    recs, in_record, have_seq, record_line = [], False, False, 0
    while i < total:
        c = data[i]
        if not in_record:             # fasta.hpp:105-137
            if c in (nl, cr, sp, tab):
                i, line = consume_eol(i, line)
                continue
            if c != gt:
                return None, "expected '>' at start of a record, found '%s' (line %d)" % (
                    chr(c), line)
            e = eol(i + 1)
            header = data[i + 1:e].rstrip(b" \t")
            i = e
            cut = [k for k, ch in enumerate(header) if ch in (sp, tab)]
            rid = header if not cut else header[:cut[0]]
            desc = b"" if not cut else header[cut[0]:].lstrip(b" \t")
            if not rid:
                return None, "empty identifier in header (line %d)" % line
            recs.append([rid, desc, []])
            in_record, have_seq, record_line = True, False, line
            i, line = consume_eol(i, line)
            continue
        if c == gt:                   # fasta.hpp:141-146
            if not have_seq:
                return None, "record has a header but no sequence (line %d)" % record_line
            in_record = False
            continue
        if c in (nl, cr):
            i, line = consume_eol(i, line)
            continue
        e = eol(i)                    # fasta.hpp:150-153
        recs[-1][2].append(data[i:e])
        i = e
        have_seq = True
    if in_record and not have_seq:
        return None, "record has a header but no sequence (line %d)" % record_line
    if not recs:
        return None, "no records found (line %d)" % line
    return [(r[0], r[1], b"".join(r[2])) for r in recs], None


def ref_fasta(data: bytes):
    """The reference's own io::read_fasta (compiled into oracle/_ref), same
    return convention as fasta_parse."""
    data = _b(data)
    R = ref()
    blen = C.c_longlong()
    err = C.create_string_buffer(512)
    n = R.ga_ref_fasta_parse(data, len(data), C.byref(blen), err, 512)
    if n == -1:
        return None, err.value.decode("latin-1")
    assert n >= 0, "ga_ref_fasta_parse: tmpfile failed"
    blob = C.create_string_buffer(max(blen.value, 1))
    R.ga_ref_fasta_blob(blob)
    raw = blob.raw[:blen.value]
    out, o = [], 0
    for _ in range(n):
        li, ld, ls = np.frombuffer(raw, dtype=np.int64, count=3, offset=o)
        o += 24
        f = []
        for ln in (li, ld, ls):
            f.append(raw[o:o + int(ln)])
            o += int(ln)
        out.append(tuple(f))
    return out, None


def ref_fasta_file(path: str):
    """(record count, None) or (None, error) from io::read_fasta_file."""
    err = C.create_string_buffer(512)
    n = ref().ga_ref_fasta_file(os.fsencode(path), err, 512)
    return (None, err.value.decode("latin-1")) if n < 0 else (int(n), None)


def fasta_pack2(seq: bytes) -> np.ndarray:
    """uint32 2-bit codes of a byte stream, 16 bases a word, base p at bits
    2(p mod 16) of word p/16: A 0, C 1, G 2, T 3, any other byte 0."""
    a = np.frombuffer(_b(seq), dtype=np.uint8)
    code = np.zeros(len(a), dtype=np.uint32)
    for ch, v in ((b"C", 1), (b"G", 2), (b"T", 3)):
        code[a == ch[0]] = v
    pad = (-len(a)) % 16
    code = np.concatenate([code, np.zeros(pad, np.uint32)]).reshape(-1, 16)
    return (code << (2 * np.arange(16, dtype=np.uint32))).sum(axis=1, dtype=np.uint64).astype(np.uint32)


# ------------------------------------------------------------------ gene QC
REF_GENE_QC = os.path.join(HERE, "_ref", "gene_qc")


def _is_stop(a: int, b: int, c: int, code: int) -> bool:
    """tools/gene_qc.cpp:53-61."""
    t = bytes((a, b, c))
    if t in (b"TAA", b"TAG"):
        return True
    if code == 1 and t == b"TGA":
        return True
    return code == 2 and t in (b"AGA", b"AGG")


def gene_qc_walk(cigar: bytes, text: bytes, score: int, resolved: bool, start_i: int, start_j: int,
                 end_i: int, frame: int, code: int):
    """(cov_text, fs_runs, fs_net, stops, stop_end) of one record: a
    restatement of tools/gene_qc.cpp:146-217."""
    cigar, text = _b(cigar), _b(text)
    if not resolved or score <= 0:
        return 0.0, 0, 0, 0, 0
    cov = (end_i - start_i + 1) / len(text)
    i, j = start_i, start_j
    codon, fill, i0, last_codon_j, stop_at, stops = [0, 0, 0], -1, -1, -1, -1, 0
    for op in cigar:
        if op in b"MX":
            pos = (j - frame) % 3
            if pos == 0:
                fill, i0 = 0, i
            if fill == pos:
                codon[fill] = text[i]
                fill += 1
            else:
                fill = -1
            if fill == 3:
                if i == i0 + 2:
                    last_codon_j = j - 2
                    if _is_stop(codon[0], codon[1], codon[2], code):
                        stops += 1
                        stop_at = j - 2
                fill = -1
            i += 1
            j += 1
        elif op == ord("I"):
            i += 1
        elif op == ord("D"):
            fill = -1
            j += 1
    runs, ins, dele = 0, 0, 0
    p = 0
    while p < len(cigar):
        o = cigar[p]
        if o not in b"ID":
            p += 1
            continue
        q = p
        while q < len(cigar) and cigar[q] == o:
            q += 1
        runs += (q - p) % 3 != 0
        if o == ord("I"):
            ins += q - p
        else:
            dele += q - p
        p = q
    net = (ins - dele) % 3
    if net > 1:
        net -= 3
    stop_end = 0
    if stops > 0 and stop_at == last_codon_j:
        stop_end, stops = 1, stops - 1
    return cov, runs, net, stops, stop_end


def gene_qc_bait(seqs, code: int):
    """(record, frame, stops) of the bait choice, tools/gene_qc.cpp:111-131."""
    best, best_stops, frame, best_atg = 0, 1 << 30, 0, False
    for r, s in enumerate(seqs):
        s = _b(s)
        for f in range(3):
            st, i = 0, f
            while i + 2 < len(s):
                st += _is_stop(s[i], s[i + 1], s[i + 2], code)
                i += 3
            atg = f + 2 < len(s) and s[f:f + 3] == b"ATG"
            if (st < best_stops or (st == best_stops and atg and not best_atg) or
                    (st == best_stops and atg == best_atg and len(s) > len(_b(seqs[best])))):
                best_stops, best, frame, best_atg = st, r, f, atg
    return best, frame, best_stops


def ref_gene_qc(seqs_path: str, ref_path: str, args=(), emit: Optional[str] = None):
    """Run the reference's gene_qc (oracle/_ref/gene_qc): (exit status, stdout,
    stderr, emitted TSV or None)."""
    cmd = [REF_GENE_QC, seqs_path, "--ref", ref_path, *args] + (["--emit", emit] if emit else [])
    p = subprocess.run(cmd, capture_output=True, timeout=600)
    tsv = open(emit, "rb").read().decode("latin-1") if emit and os.path.exists(emit) else None
    return p.returncode, p.stdout.decode("latin-1"), p.stderr.decode("latin-1"), tsv


def gene_qc_tsv(seqs: bytes, refs: bytes, code: int = 1, scoring=(1, -1, 2, 1), min_cov: float = 0.5,
                min_score: int = 0):
    """The reference gene_qc tool's TSV (tools/gene_qc.cpp:175-244) from the
    oracle's FASTA reader, SW batch and walk: (tsv text, summary line)."""
    recs, err = fasta_parse(seqs)
    rrecs, err2 = fasta_parse(refs)
    assert err is None and err2 is None, (err, err2)
    best, frame, _ = gene_qc_bait([r[2] for r in rrecs], code)
    bait = rrecs[best][2]

    @dataclass
    class _Req:
        text: bytes
        text_len: int
        pattern: bytes
        pattern_len: int
        scoring: tuple
        with_cigar: bool = True
    br = align_sw_batch([_Req(r[2], len(r[2]), bait, len(bait), tuple(scoring)) for r in recs])
    assert br.status == "ok", br.error
    lines = ["acc\tlen\tscore\tsi\tsj\tei\tej\tcov_text\tfs_runs\tfs_net\tstops\tstop_end\tflag"]
    cnt = dict(ok=0, na=0, lc=0, fs=0, st=0, bad=0)
    for rec, r in zip(recs, br.results):
        cov, runs, net, stops, stop_end = gene_qc_walk(r.cigar, rec[2], r.score, r.resolved, r.start_i,
                                                       r.start_j, r.end_i, frame, code)
        if not r.resolved or r.score <= min_score:
            flag = "NO_ALIGN"
        else:
            flag = "LOW_COV" if cov < min_cov else ""
            if net != 0:
                flag += ("+" if flag else "") + "FS=%d" % net
            if stops > 0:
                flag += ("+" if flag else "") + "STOP=%d" % stops
            flag = flag or "OK"
        lines.append("%s\t%d\t%d\t%d\t%d\t%d\t%d\t%.3f\t%d\t%d\t%d\t%d\t%s" % (
            rec[0].decode("latin-1"), len(rec[2]), r.score, r.start_i, r.start_j, r.end_i, r.end_j, cov, runs,
            net, stops, stop_end, flag))
        if flag == "OK":
            cnt["ok"] += 1
        else:
            cnt["na"] += "NO_ALIGN" in flag
            cnt["lc"] += "LOW_COV" in flag
            cnt["fs"] += "FS=" in flag
            cnt["st"] += "STOP=" in flag
        cnt["bad"] += r.resolved and r.score > 0 and (not r.rescore_ok or not r.wellformed_ok)
    summary = "summary: %d seqs | OK=%d NO_ALIGN=%d LOW_COV=%d FS=%d STOP=%d | invalid=%d" % (
        len(recs), cnt["ok"], cnt["na"], cnt["lc"], cnt["fs"], cnt["st"], cnt["bad"])
    return "\n".join(lines) + "\n", summary


# ---------------------------------------------------------------------------
# MSA building blocks (src/msa/msa_ref.cpp): profiles, the profile-profile
# Gotoh, k-mer distances, neighbour joining
# ---------------------------------------------------------------------------
REF_MSA_SO = os.path.join(HERE, "_ref", "libga_ref_msa.so")

# msa_ref.cpp:117-138 (rows and columns in ARNDCQEGHILKMFPSTWYV order)
MSA_BLOSUM62 = np.array([
    [4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0],
    [-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3],
    [-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3],
    [-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3],
    [0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1],
    [-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2],
    [-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2],
    [0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3],
    [-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3],
    [-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3],
    [-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1],
    [-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2],
    [-1, -1, -2, -3, -1, 0, -2, -3, -1, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1],
    [-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1],
    [-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2],
    [1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2],
    [0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0],
    [-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3],
    [-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1],
    [0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4],
], dtype=np.float32)


class MsaParamsC(C.Structure):
    _fields_ = [("match", C.c_float), ("ts", C.c_float), ("tv", C.c_float),
                ("gap_open", C.c_float), ("gap_extend", C.c_float), ("free_end_gaps", C.c_int),
                ("psgp", C.c_int), ("psgp_scale", C.c_float), ("psgp_min_open", C.c_float),
                ("psgp_min_ext", C.c_float), ("alpha", C.c_int), ("sub", C.c_void_p)]


@dataclass
class MsaParams:
    """genomsa::Params fields of the profile DP (msa.hpp:36-60 defaults);
    protein() is protein_params() (msa_ref.cpp:140-151)."""
    match: float = 2.0
    ts: float = -1.0
    tv: float = -2.0
    gap_open: float = 3.0
    gap_extend: float = 1.0
    free_end_gaps: bool = True
    psgp: bool = True
    psgp_scale: float = 0.5
    psgp_min_open: float = 0.1
    psgp_min_ext: float = 0.2
    alpha: int = 4
    sub: Optional[np.ndarray] = None

    @staticmethod
    def protein() -> "MsaParams":
        return MsaParams(gap_open=11.0, gap_extend=1.0, free_end_gaps=False, alpha=20,
                         sub=MSA_BLOSUM62.copy())

    def c(self):
        sub = None if self.sub is None else np.ascontiguousarray(self.sub, dtype=np.float32)
        p = MsaParamsC(self.match, self.ts, self.tv, self.gap_open, self.gap_extend, int(self.free_end_gaps),
                       int(self.psgp), self.psgp_scale, self.psgp_min_open, self.psgp_min_ext, self.alpha,
                       sub.ctypes.data if sub is not None else None)
        p._keep = sub
        return p


def _msa_lib():
    L = lib()
    if not getattr(L, "_msa_ready", False):
        vp = C.c_void_p
        L.ga_oracle_msa_profile.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp]
        L.ga_oracle_msa_align.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int, C.POINTER(MsaParamsC), vp, vp, vp, vp]
        L.ga_oracle_msa_align_batch.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int64, C.POINTER(MsaParamsC),
                                                vp, vp, vp, vp, vp, C.c_int]
        L.ga_oracle_kmer_distances.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp]
        L.ga_oracle_nj_tree.argtypes = [vp, C.c_int, vp]
        L._msa_ready = True
    return L


def msa_profile(rows: Sequence, alpha: int = 4):
    """(cols float32 [L, alpha], occ float32 [L]) of aligned rows."""
    rows = [_b(r) for r in rows]
    L = len(rows[0]) if rows else 0
    assert all(len(r) == L for r in rows)
    cols = np.zeros((L, alpha), dtype=np.float32)
    occ = np.zeros(L, dtype=np.float32)
    rc = _msa_lib().ga_oracle_msa_profile(b"".join(rows), len(rows), L, alpha, cols.ctypes.data, occ.ctypes.data)
    assert rc == 0
    return cols, occ


def msa_align(A, B, params: MsaParams):
    """align_profiles: (score, (ai, aj, bi, bj), cigar) of profiles A = (cols,
    occ) and B."""
    ca, oa = (np.ascontiguousarray(x, dtype=np.float32) for x in A)
    cb, ob = (np.ascontiguousarray(x, dtype=np.float32) for x in B)
    M, N = len(oa), len(ob)
    score = np.zeros(1, dtype=np.float32)
    span = np.zeros(4, dtype=np.int32)
    cig = C.create_string_buffer(M + N + 1)
    clen = np.zeros(1, dtype=np.int32)
    p = params.c()
    rc = _msa_lib().ga_oracle_msa_align(ca.ctypes.data, oa.ctypes.data, M, cb.ctypes.data, ob.ctypes.data, N,
                                        C.byref(p), score.ctypes.data, span.ctypes.data, cig, clen.ctypes.data)
    assert rc == 0
    return float(score[0]), tuple(int(x) for x in span), cig.raw[: int(clen[0])].decode()


def msa_merge_rows(rowsA: Sequence, rowsB: Sequence, span, cigar: str):
    """The rows of merge_profiles (msa_ref.cpp:1177-1220): A's and B's rows
    interleaved by the alignment, prefix / suffix columns as gap blocks."""
    ai, aj, bi, bj = span
    la, lb = len(rowsA[0]), len(rowsB[0])
    w = ai + aj + len(cigar) + (la - bi) + (lb - bj)
    out = []
    for row in rowsA:
        s, ca = [row[:ai], "-" * aj], ai
        for op in cigar:
            if op in "MI":
                s.append(row[ca]); ca += 1
            else:
                s.append("-")
        s += [row[ca:], "-" * (lb - bj)]
        out.append("".join(s).ljust(w, "-"))
    for row in rowsB:
        s, cb = ["-" * ai, row[:aj]], aj
        for op in cigar:
            if op in "MD":
                s.append(row[cb]); cb += 1
            else:
                s.append("-")
        s += ["-" * (la - bi), row[cb:]]
        out.append("".join(s).ljust(w, "-"))
    return out


def kmer_distances(seqs: Sequence, k: int, alpha: int = 4) -> np.ndarray:
    """Packed upper-triangle k-mer distances (kmer_distances)."""
    bs = [_b(s) for s in seqs]
    lens = np.array([len(s) for s in bs], dtype=np.int32)
    off = np.zeros(len(bs), dtype=np.int64)
    off[1:] = np.cumsum(lens.astype(np.int64))[:-1] if len(bs) > 1 else off[1:]
    buf = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8)
    n = len(bs)
    D = np.zeros(max(n * (n - 1) // 2, 1), dtype=np.float32)
    rc = _msa_lib().ga_oracle_kmer_distances(buf.ctypes.data, off.ctypes.data, lens.ctypes.data, n, k, alpha,
                                             D.ctypes.data)
    assert rc == 0
    return D[: n * (n - 1) // 2]


def nj_tree(D: np.ndarray, n: int) -> np.ndarray:
    """nj_tree: int32 [n-1, 2] children of internal nodes n .. 2n-2."""
    D = np.ascontiguousarray(D, dtype=np.float32)
    ch = np.zeros(2 * (n - 1), dtype=np.int32)
    rc = _msa_lib().ga_oracle_nj_tree(D.ctypes.data, n, ch.ctypes.data)
    assert rc == 0
    return ch.reshape(n - 1, 2)


_ref_msa = None


def ref_msa():
    """The reference's MSA code compiled for the CPU (None when not built)."""
    global _ref_msa
    if _ref_msa is None and os.path.exists(REF_MSA_SO):
        R = C.CDLL(REF_MSA_SO)
        vp, cp = C.c_void_p, C.c_char_p
        R.ga_ref_msa_profile.argtypes = [cp, C.c_int, C.c_int, C.c_int, vp, vp]
        for f in (R.ga_ref_msa_align, R.ga_ref_msa_pp_kernel):
            f.argtypes = [cp, C.c_int, C.c_int, cp, C.c_int, C.c_int, C.POINTER(MsaParamsC), vp, vp, vp, vp]
        R.ga_ref_kmer_distances.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp]
        R.ga_ref_nj_tree.argtypes = [vp, C.c_int, vp]
        _ref_msa = R
    return _ref_msa


def ref_msa_profile(rows: Sequence, alpha: int = 4):
    rows = [_b(r) for r in rows]
    L = len(rows[0])
    cols = np.zeros((L, alpha), dtype=np.float32)
    occ = np.zeros(L, dtype=np.float32)
    ref_msa().ga_ref_msa_profile(b"".join(rows), len(rows), L, alpha, cols.ctypes.data, occ.ctypes.data)
    return cols, occ


def ref_msa_align(rowsA: Sequence, rowsB: Sequence, params: MsaParams, kernel: bool = False):
    """align_profiles (or, with kernel, the shipped msa_pp_trace_kernel body)
    on the profiles of two row sets."""
    ra, rb = [_b(r) for r in rowsA], [_b(r) for r in rowsB]
    M, N = len(ra[0]), len(rb[0])
    score = np.zeros(1, dtype=np.float32)
    span = np.zeros(4, dtype=np.int32)
    cig = C.create_string_buffer(M + N + 8)
    clen = np.zeros(1, dtype=np.int32)
    p = params.c()
    f = ref_msa().ga_ref_msa_pp_kernel if kernel else ref_msa().ga_ref_msa_align
    f(b"".join(ra), len(ra), M, b"".join(rb), len(rb), N, C.byref(p), score.ctypes.data, span.ctypes.data, cig,
      clen.ctypes.data)
    return float(score[0]), tuple(int(x) for x in span), cig.raw[: int(clen[0])].decode()


def ref_kmer_distances(seqs: Sequence, k: int, alpha: int = 4) -> np.ndarray:
    bs = [_b(s) for s in seqs]
    lens = np.array([len(s) for s in bs], dtype=np.int32)
    off = np.zeros(len(bs), dtype=np.int64)
    if len(bs) > 1:
        off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    buf = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8)
    n = len(bs)
    D = np.zeros(max(n * (n - 1) // 2, 1), dtype=np.float32)
    ref_msa().ga_ref_kmer_distances(buf.ctypes.data, off.ctypes.data, lens.ctypes.data, n, k, alpha, D.ctypes.data)
    return D[: n * (n - 1) // 2]


def ref_nj_tree(D: np.ndarray, n: int) -> np.ndarray:
    D = np.ascontiguousarray(D, dtype=np.float32)
    ch = np.zeros(2 * (n - 1), dtype=np.int32)
    ref_msa().ga_ref_nj_tree(D.ctypes.data, n, ch.ctypes.data)
    return ch.reshape(n - 1, 2)
