"""The MSA building blocks of the reference's genomsa engine on the GPU:
profiles, the profile-profile Gotoh alignment (align_profiles), and (for
the guide tree) k-mer distances and neighbour joining.

Mirrors include/genoaligner/msa/msa.hpp of the reference: ``Params`` holds
the fields of genomsa::Params the DP reads (same defaults; ``protein_params``
is genomsa::protein_params), ``Profile`` the columns' symbol fractions and
occupancy of a set of aligned rows (genomsa::profile_from_seq /
profile_update_counts), ``align_profiles`` returns genomsa::AlignResult's
fields.  Everything runs through the C ABI of libgenoaligner_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

# BLOSUM62 as the reference's MSA engine holds it (src/msa/msa_ref.cpp:
# 117-138), rows and columns in ARNDCQEGHILKMFPSTWYV order.
_B62 = (
    "4 -1 -2 -2 0 -1 -1 0 -2 -1 -1 -1 -1 -2 -1 1 0 -3 -2 0 "
    "-1 5 0 -2 -3 1 0 -2 0 -3 -2 2 -1 -3 -2 -1 -1 -3 -2 -3 "
    "-2 0 6 1 -3 0 0 0 1 -3 -3 0 -2 -3 -2 1 0 -4 -2 -3 "
    "-2 -2 1 6 -3 0 2 -1 -1 -3 -4 -1 -3 -3 -1 0 -1 -4 -3 -3 "
    "0 -3 -3 -3 9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1 "
    "-1 1 0 0 -3 5 2 -2 0 -3 -2 1 0 -3 -1 0 -1 -2 -1 -2 "
    "-1 0 0 2 -4 2 5 -2 0 -3 -3 1 -2 -3 -1 0 -1 -3 -2 -2 "
    "0 -2 0 -1 -3 -2 -2 6 -2 -4 -4 -2 -3 -3 -2 0 -2 -2 -3 -3 "
    "-2 0 1 -1 -3 0 0 -2 8 -3 -3 -1 -2 -1 -2 -1 -2 -2 2 -3 "
    "-1 -3 -3 -3 -1 -3 -3 -4 -3 4 2 -3 1 0 -3 -2 -1 -3 -1 3 "
    "-1 -2 -3 -4 -1 -2 -3 -4 -3 2 4 -2 2 0 -3 -2 -1 -2 -1 1 "
    "-1 2 0 -1 -3 1 1 -2 -1 -3 -2 5 -1 -3 -1 0 -1 -3 -2 -2 "
    "-1 -1 -2 -3 -1 0 -2 -3 -1 1 2 -1 5 0 -2 -1 -1 -1 -1 1 "
    "-2 -3 -3 -3 -2 -3 -3 -3 -1 0 0 -3 0 6 -4 -2 -2 1 3 -1 "
    "-1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4 7 -1 -1 -4 -3 -2 "
    "1 -1 1 0 -1 0 0 0 -1 -2 -2 0 -1 -2 -1 4 1 -3 -2 -2 "
    "0 -1 0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1 1 5 -2 -2 0 "
    "-3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1 1 -4 -3 -2 11 2 -3 "
    "-2 -2 -2 -3 -2 -1 -2 -3 2 -1 -1 -2 -1 3 -3 -2 -2 2 7 -1 "
    "0 -3 -3 -3 -1 -2 -2 -3 -3 3 1 -2 1 -1 -2 -2 0 -3 -1 4")
BLOSUM62 = np.array([float(x) for x in _B62.split()], dtype=np.float32).reshape(20, 20)


class MsaError(RuntimeError):
    pass


@dataclass
class Params:
    """genomsa::Params (msa.hpp:36-60): the fields the profile DP and the
    guide tree read, with the reference's defaults (DNA, semiglobal,
    position-specific gap penalties on)."""
    kmer_k: int = 5
    alpha: int = 4
    match: float = 2.0
    ts: float = -1.0
    tv: float = -2.0
    sub: Optional[np.ndarray] = None        # alpha x alpha, alpha 20
    gap_open: float = 3.0
    gap_extend: float = 1.0
    free_end_gaps: bool = True
    psgp: bool = True
    psgp_scale: float = 0.5
    psgp_min_open: float = 0.1
    psgp_min_ext: float = 0.2

    def c(self) -> N.MsaParams:
        p = N.MsaParams(self.match, self.ts, self.tv, self.gap_open, self.gap_extend, int(self.free_end_gaps),
                        int(self.psgp), self.psgp_scale, self.psgp_min_open, self.psgp_min_ext, self.alpha)
        if self.alpha == 20:
            sub = BLOSUM62 if self.sub is None else np.asarray(self.sub, dtype=np.float32).reshape(20, 20)
            for i, v in enumerate(sub.ravel().tolist()):
                p.sub[i] = v
        return p


def protein_params() -> Params:
    """genomsa::protein_params (msa_ref.cpp:140-151)."""
    return Params(kmer_k=2, alpha=20, sub=BLOSUM62.copy(), gap_open=11.0, gap_extend=1.0, free_end_gaps=False)


@dataclass
class Profile:
    """Columns of symbol fractions (cols[c, s]) and occupancy (occ[c]) of a
    set of aligned rows (genomsa::Profile without the gap-fraction slot,
    which the DP never reads)."""
    cols: np.ndarray
    occ: np.ndarray
    rows: List[str] = field(default_factory=list)
    alpha: int = 4

    def ncols(self) -> int:
        return int(self.occ.shape[0])


def profile_from_rows(rows: Sequence[str], alpha: int = 4) -> Profile:
    """genomsa::profile_update_counts over aligned rows of equal length."""
    rows = [r if isinstance(r, str) else bytes(r).decode("latin-1") for r in rows]
    if not rows or len({len(r) for r in rows}) != 1:
        raise MsaError("profile rows must be non-empty and of equal length")
    L = len(rows[0])
    cols = np.zeros((L, alpha), dtype=np.float32)
    occ = np.zeros(L, dtype=np.float32)
    rc = N.lib().ga_msa_profile_from_rows("".join(rows).encode("latin-1"), len(rows), L, alpha,
                                          cols.ctypes.data, occ.ctypes.data, None)
    if rc != N.GA_OK:
        raise MsaError("invalid profile rows or alphabet")
    return Profile(cols, occ, list(rows), alpha)


def profile_from_seq(seq: str, alpha: int = 4) -> Profile:
    """genomsa::profile_from_seq (msa_ref.cpp:751-771)."""
    return profile_from_rows([seq], alpha)


@dataclass
class AlignResult:
    """genomsa::AlignResult (msa.hpp:339-344)."""
    score: float = 0.0
    cigar: str = ""
    ai: int = 0
    aj: int = 0
    bi: int = 0
    bj: int = 0


def align_profiles_batch(pairs: Sequence[Tuple[Profile, Profile]], P: Params) -> List[AlignResult]:
    """genomsa::align_profiles for many independent pairs in one launch (the
    reference's level-batched msa_pp_trace_kernel), bit-exact floats."""
    n = len(pairs)
    if n == 0:
        return []
    A = (N.MsaProfile * n)()
    B = (N.MsaProfile * n)()
    keep = []
    for q, (a, b) in enumerate(pairs):
        for arr, prof in ((A, a), (B, b)):
            cols = np.ascontiguousarray(prof.cols, dtype=np.float32)
            occ = np.ascontiguousarray(prof.occ, dtype=np.float32)
            keep += [cols, occ]
            arr[q] = N.MsaProfile(cols.ctypes.data if cols.size else None, occ.ctypes.data if occ.size else None,
                                  int(occ.shape[0]))
    res = (N.MsaPPRes * n)()
    need = sum(a.ncols() + b.ncols() for a, b in pairs)
    arena = C.create_string_buffer(max(need, 1))
    err = C.c_char_p()
    p = P.c()
    rc = N.lib().ga_msa_align_profiles(A, B, n, C.byref(p), res, arena, need, C.byref(err))
    if rc != N.GA_OK:
        raise MsaError((err.value or b"MSA alignment failed").decode())
    raw = arena.raw
    return [AlignResult(r.score, raw[r.cigar_off: r.cigar_off + r.cigar_len].decode(), r.ai, r.aj, r.bi, r.bj)
            for r in res]


def align_profiles(A: Profile, B: Profile, P: Params) -> AlignResult:
    """genomsa::align_profiles (msa.hpp:345)."""
    return align_profiles_batch([(A, B)], P)[0]


def merge_rows(A: Profile, B: Profile, aln: AlignResult) -> List[str]:
    """The rows of genomsa::merge_profiles (msa_ref.cpp:1177-1220): both
    profiles' rows interleaved by the alignment, the columns outside the
    span as gap blocks."""
    la, lb = A.ncols(), B.ncols()
    out = []
    for row in A.rows:
        s, ca = [row[: aln.ai], "-" * aln.aj], aln.ai
        for op in aln.cigar:
            if op in "MI":
                s.append(row[ca]); ca += 1
            else:
                s.append("-")
        s += [row[ca:], "-" * (lb - aln.bj)]
        out.append("".join(s))
    for row in B.rows:
        s, cb = ["-" * aln.ai, row[: aln.aj]], aln.aj
        for op in aln.cigar:
            if op in "MD":
                s.append(row[cb]); cb += 1
            else:
                s.append("-")
        s += ["-" * (la - aln.bi), row[cb:]]
        out.append("".join(s))
    return out


def merge_profiles(A: Profile, B: Profile, aln: AlignResult) -> Profile:
    """genomsa::merge_profiles: the merged rows and their profile."""
    return profile_from_rows(merge_rows(A, B, aln), A.alpha)


class MsaDeviceBatch:
    """Profiles and pairs resident on the device (ga_msa_pp_device): the
    profiles' columns in one table, pair q aligning profiles[pairs[q][0]]
    against profiles[pairs[q][1]]."""

    def __init__(self, profiles: Sequence[Profile], pairs: Sequence[Tuple[int, int]], P: Params,
                 device: str = "cuda:0"):
        import torch
        dev = torch.device(device)
        self.P = P
        self.p = P.c()
        self.n = len(pairs)
        lens = np.array([p.ncols() for p in profiles], dtype=np.int32)
        off = np.zeros(len(profiles), dtype=np.int64)
        if len(profiles) > 1:
            off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
        al = P.alpha
        cols = np.concatenate([np.asarray(p.cols, dtype=np.float32).reshape(-1, al) for p in profiles]) \
            if int(lens.sum()) else np.zeros((1, al), np.float32)
        occ = np.concatenate([np.asarray(p.occ, dtype=np.float32) for p in profiles]) \
            if int(lens.sum()) else np.zeros(1, np.float32)
        self.h_len = lens
        self.h_a = np.ascontiguousarray([a for a, _ in pairs], dtype=np.int32)
        self.h_b = np.ascontiguousarray([b for _, b in pairs], dtype=np.int32)
        slot = lens[self.h_a].astype(np.int64) + lens[self.h_b]
        self.slot = slot
        self.slot_off = np.zeros(self.n, dtype=np.int64)
        if self.n > 1:
            self.slot_off[1:] = np.cumsum(slot)[:-1]
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.cols, self.occ, self.col_off, self.len = t(cols), t(occ), t(off), t(lens)
        self.a_idx, self.b_idx, self.cig_off = t(self.h_a), t(self.h_b), t(self.slot_off)
        self.score = torch.empty(max(self.n, 1), dtype=torch.float32, device=dev)
        self.span = torch.empty(max(self.n, 1) * 4, dtype=torch.int32, device=dev)
        self.cig_len = torch.empty(max(self.n, 1), dtype=torch.int32, device=dev)
        self.cigar = torch.empty(max(int(slot.sum()), 1), dtype=torch.uint8, device=dev)
        self.b = N.MsaDevBatch(self.cols.data_ptr(), self.occ.data_ptr(), int(occ.shape[0]), self.col_off.data_ptr(),
                               self.len.data_ptr(), len(profiles), self.a_idx.data_ptr(), self.b_idx.data_ptr(),
                               self.n, self.score.data_ptr(), self.span.data_ptr(), self.cigar.data_ptr(),
                               self.cig_off.data_ptr(), self.cig_len.data_ptr(), self.h_len.ctypes.data,
                               self.h_a.ctypes.data, self.h_b.ctypes.data)

    def run(self, stream=None):
        from .device import _check, _stream_handle
        err = C.c_char_p()
        _check(N.lib().ga_msa_pp_device(C.byref(self.b), C.byref(self.p), C.c_void_p(_stream_handle(stream)),
                                        C.byref(err)), err, "ga_msa_pp_device")

    def results(self) -> List[AlignResult]:
        import torch
        torch.cuda.synchronize()
        sc = self.score.cpu().numpy()
        sp = self.span.cpu().numpy().reshape(-1, 4)
        ln = self.cig_len.cpu().numpy()
        raw = self.cigar.cpu().numpy().tobytes()
        out = []
        for q in range(self.n):
            end = int(self.slot_off[q] + self.slot[q])
            out.append(AlignResult(float(sc[q]), raw[end - int(ln[q]): end].decode(), *(int(x) for x in sp[q])))
        return out


# ---------------------------------------------------------------- guide tree
@dataclass
class Tree:
    """genomsa::Tree (msa.hpp:338-341): nodes[0 .. n-1] are the leaves
    (-1, -1), nodes[n ..] the internal nodes' (left, right) children."""
    nodes: List[Tuple[int, int]]
    root: int


def kmer_distances(seqs: Sequence, k: int = 5, alpha: int = 4) -> np.ndarray:
    """genomsa::kmer_distances: packed upper-triangle float32 distances."""
    bs = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in seqs]
    n = len(bs)
    if n < 2:
        return np.zeros(0, dtype=np.float32)
    lens = np.array([len(s) for s in bs], dtype=np.int32)
    off = np.zeros(n, dtype=np.int64)
    off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    buf = np.frombuffer(b"".join(bs) + b"\0", dtype=np.uint8)
    D = np.zeros(n * (n - 1) // 2, dtype=np.float32)
    err = C.c_char_p()
    rc = N.lib().ga_kmer_distances(buf.ctypes.data, off.ctypes.data, lens.ctypes.data, n, k, alpha, D.ctypes.data,
                                   C.byref(err))
    if rc != N.GA_OK:
        raise MsaError((err.value or b"k-mer distances failed").decode())
    return D


def nj_children(D: np.ndarray, n: int) -> np.ndarray:
    """ga_nj_tree: int32 [n-1, 2] children of internal nodes n .. 2n-2."""
    if n < 2:
        return np.zeros((0, 2), dtype=np.int32)
    D = np.ascontiguousarray(D, dtype=np.float32)
    if D.shape[0] != n * (n - 1) // 2:
        raise MsaError("distance array is not n (n - 1) / 2 long")
    ch = np.zeros(2 * (n - 1), dtype=np.int32)
    err = C.c_char_p()
    rc = N.lib().ga_nj_tree(D.ctypes.data, n, ch.ctypes.data, C.byref(err))
    if rc != N.GA_OK:
        raise MsaError((err.value or b"neighbour joining failed").decode())
    return ch.reshape(n - 1, 2)


def nj_tree(D: np.ndarray, n: int) -> Tree:
    """genomsa::nj_tree (msa_ref.cpp:904-947) on the device."""
    if n == 1:
        return Tree([(-1, -1)], 0)
    ch = nj_children(D, n)
    return Tree([(-1, -1)] * n + [tuple(int(x) for x in c) for c in ch], 2 * n - 2)
