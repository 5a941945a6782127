"""Gene QC tool (reference: tools/gene_qc.cpp): every record of a gene
FASTA aligned to the gene's bait by Smith-Waterman with CIGAR on the GPU (one
pattern, many texts), then each CIGAR walked for coverage, frameshifts and
in-frame stops (``ga_gene_qc_walk``, native, on the host's worker threads),
and the TSV written by ``ga_gene_qc_tsv`` (native too).  The records go from
the FASTA reader's buffers to the batch with no Python object per record.

Same command line, the same TSV rows and the same summary as the reference
tool: ``python -m genoaligner_amd.gene_qc seqs.fa --ref ref.fa [--code 1|2]
[--emit out.tsv] [--scoring m,x,go,ge] [--min-cov f] [--min-score n]
[--batch-size n]``.  The one difference is the batch size's default: 0, all
records in one batch (the reference's 400 sized its own workspace); a batch
the library refuses is halved and tried again, as there.
"""
from __future__ import annotations

import ctypes as C
import sys
from typing import List, Optional, Sequence, TextIO, Tuple

import numpy as np

from . import _native as N
from .api import _SW_REQ, _SW_RES, _ptr, backend_name
from .fasta import FastaError, read_fasta_file, read_fasta_file_arrays

# tools/gene_qc.cpp:53-61, by genetic code (any other code: TAA and TAG only)
_STOPS = {1: (b"TAA", b"TAG", b"TGA"), 2: (b"TAA", b"TAG", b"AGA", b"AGG")}
_STOPS_OTHER = (b"TAA", b"TAG")


def stops_in_frame(seq: bytes, frame: int, code: int) -> int:
    """Stops among the codons at frame, frame+3, ... (gene_qc.cpp:63-69)."""
    n = (len(seq) - frame) // 3 if len(seq) - frame >= 3 else 0
    if n <= 0:
        return 0
    c = np.frombuffer(seq, np.uint8, count=3 * n, offset=frame).reshape(n, 3)
    hit = np.zeros(n, bool)
    for s in _STOPS.get(code, _STOPS_OTHER):
        hit |= (c[:, 0] == s[0]) & (c[:, 1] == s[1]) & (c[:, 2] == s[2])
    return int(hit.sum())


def pick_bait(seqs: Sequence[bytes], code: int) -> Tuple[int, int, int]:
    """(record, frame, stops): the fewest stops; ties to a frame starting ATG,
    then to the longer record (gene_qc.cpp:111-131)."""
    best, best_stops, frame, best_atg = 0, 1 << 30, 0, False
    for r, s in enumerate(seqs):
        for f in range(3):
            st = stops_in_frame(s, f, code)
            atg = f + 2 < len(s) and s[f:f + 3] == b"ATG"
            better = (st < best_stops or (st == best_stops and atg and not best_atg) or
                      (st == best_stops and atg == best_atg and len(s) > len(seqs[best])))
            if better:
                best_stops, best, frame, best_atg = st, r, f, atg
    return best, frame, best_stops


def _align(buf: np.ndarray, off: np.ndarray, lens: np.ndarray, bait: bytes, frame: int, code: int, scoring,
           batch_size: int, err: TextIO):
    """align_sw_batch in chunks (halved on a refusal) + the native walk over
    the records buf[off[k]:][:lens[k]].  Returns (res structured array, qc rows
    array) or None after a failure."""
    L = N.lib()
    n = len(lens)
    pat = np.frombuffer(bait + b"\0", np.uint8)
    res = np.zeros(n, dtype=_SW_RES)
    rows = np.zeros(n, dtype=np.dtype([("cov_text", "<f8"), ("fs_runs", "<i4"), ("fs_net", "<i4"),
                                      ("stops", "<i4"), ("stop_end", "<i4")]))
    assert rows.itemsize == C.sizeof(N.QcRow)
    bs = n if batch_size <= 0 else batch_size
    k = 0
    while k < n:
        m = min(bs, n - k)
        req = np.zeros(m, dtype=_SW_REQ)
        req["text"] = buf.ctypes.data + off[k:k + m]
        req["text_len"] = lens[k:k + m]
        req["pattern"] = pat.ctypes.data
        req["pattern_len"] = len(bait)
        req["scoring"] = tuple(scoring)
        req["with_cigar"] = 1
        r = np.zeros(m, dtype=_SW_RES)
        pq, pr = _ptr(req, N.SWReq), _ptr(r, N.SWRes)
        need = L.ga_cigar_arena_bytes_sw(pq, m)
        arena = np.empty(max(need, 1), np.uint8)
        ok, un, e = C.c_int32(), C.c_int32(), C.c_char_p()
        rc = L.ga_align_sw_batch(pq, m, pr, arena.ctypes.data, need, C.byref(ok), C.byref(un), C.byref(e))
        if rc != N.GA_OK:
            msg = (e.value or b"").decode()
            if m == 1:
                print(f"batch: {msg}", file=err)
                return None
            bs = max(1, bs // 2)
            print(f"batch of {m} too big ({msg}); retrying at {bs}", file=err)
            continue
        rc = L.ga_gene_qc_walk(r.ctypes.data, m, arena.ctypes.data, buf.ctypes.data, off[k:k + m].ctypes.data,
                               lens[k:k + m].ctypes.data, frame, code,
                               rows[k:k + m].ctypes.data)
        if rc != N.GA_OK:
            raise RuntimeError("ga_gene_qc_walk refused its arguments")
        res[k:k + m] = r
        k += m
    return res, rows


def run(seqs_path: str, ref_path: str, code: int = 1, scoring=(1, -1, 2, 1), min_cov: float = 0.5,
        min_score: int = 0, batch_size: int = 0, emit: Optional[str] = None,
        out: Optional[TextIO] = None, err: Optional[TextIO] = None) -> int:
    """The tool (gene_qc.cpp:71-267); returns its exit status.  out / err
    default to the process's stdout / stderr at the time of the call."""
    out = sys.stdout if out is None else out
    err = sys.stderr if err is None else err
    try:
        recs = read_fasta_file_arrays(seqs_path)
        rrecs = read_fasta_file(ref_path)
    except FastaError as e:
        print(f"fasta error: {e}", file=err)
        return 1
    best, frame, best_stops = pick_bait([r.sequence for r in rrecs], code)
    ref = rrecs[best].sequence
    n = len(recs)
    print(f"gene_qc: {n} records vs ref {rrecs[best].id.decode('latin-1')} ({len(ref)} bp, frame "
          f"{frame}, bait_stops {best_stops}, code {code}) on {backend_name()}", file=out)
    if best_stops > 0:
        print(f"WARN: bait itself carries {best_stops} stops in its best frame -- stop counts for this gene "
              "are weak evidence", file=err)
    # the records' sequences straight from the reader's buffer
    off = np.ascontiguousarray(recs.seq_off[:n])
    lens = np.diff(recs.seq_off).astype(np.int32)
    got = _align(recs.seq, off, lens, ref, frame, code, scoring, batch_size, err)
    if got is None:
        return 1
    res, rows = got
    try:
        dst = open(emit, "wb") if emit else None
    except OSError:
        print(f"cannot write {emit}", file=err)
        return 1
    L = N.lib()
    counts = np.zeros(6, np.int64)
    args = (res.ctypes.data, rows.ctypes.data, n, recs.ids.ctypes.data, recs.id_off.ctypes.data,
            lens.ctypes.data, float(min_cov), int(min_score))
    need = L.ga_gene_qc_tsv(*args, None, 0, counts.ctypes.data)
    tsv = np.empty(max(need, 1), np.uint8)
    if need < 0 or L.ga_gene_qc_tsv(*args, tsv.ctypes.data, need, counts.ctypes.data) != need:
        raise RuntimeError("ga_gene_qc_tsv refused its arguments")
    if dst is not None:
        with dst:
            dst.write(memoryview(tsv)[:need])
    else:
        out.write(tsv[:need].tobytes().decode("latin-1"))
    n_ok, n_noalign, n_lowcov, n_fs, n_stop, n_bad = (int(x) for x in counts)
    print(f"summary: {n} seqs | OK={n_ok} NO_ALIGN={n_noalign} LOW_COV={n_lowcov} FS={n_fs} "
          f"STOP={n_stop} | invalid={n_bad}", file=err)
    return 1 if n_bad else 0


def main(argv: Optional[List[str]] = None) -> int:
    """Command line of tools/gene_qc.cpp:71-100."""
    argv = sys.argv if argv is None else argv
    if len(argv) < 4:
        print(f"usage: {argv[0]} <seqs.fasta> --ref <ref.fasta> [--code 1|2] [--emit tsv]\n"
              "       [--scoring m,x,go,ge] [--min-cov f] [--min-score n]", file=sys.stderr)
        return 2
    kw = {"code": 1, "min_cov": 0.5, "min_score": 0, "batch_size": 0, "scoring": [1, -1, 2, 1]}
    ref = emit = None
    a = 2
    while a < len(argv):
        f, v = argv[a], argv[a + 1] if a + 1 < len(argv) else None
        if v is not None and f in ("--ref", "--emit", "--code", "--min-cov", "--min-score", "--scoring",
                                   "--batch-size"):
            a += 1
            if f == "--ref":
                ref = v
            elif f == "--emit":
                emit = v
            elif f == "--code":
                kw["code"] = int(v)
            elif f == "--min-cov":
                kw["min_cov"] = float(v)
            elif f == "--min-score":
                kw["min_score"] = int(v)
            elif f == "--batch-size":
                kw["batch_size"] = int(v)
            else:
                for j, x in enumerate(v.split(",")[:4]):
                    kw["scoring"][j] = int(x)
        a += 1
    if ref is None:
        print("--ref required", file=sys.stderr)
        return 2
    return run(argv[1], ref, emit=emit, **kw)


if __name__ == "__main__":
    sys.exit(main())
