"""Seeded synthetic workloads for the pairwise path (vectorised numpy).

Each generator returns a ``PairSet``: two packed byte arrays with int64 offsets
and int32 lengths, plus per-pair metadata.  All sequences are uppercase ACGT
bytes (uint8).

Workloads (sizes as used by the benchmark and tests):
  wfa_plain   n pairs; pattern L bp i.i.d. uniform ACGT; text = pattern with
              per-base edit probability p, sub:ins:del = 2:1:1 (a substitution
              picks one of the 3 other bases, an insertion adds a uniform base
              after the current one); smax per pair.
  sw_embedded n pairs; query (pattern, columns) Lq bp i.i.d.; target (text,
              rows) Lt bp i.i.d. with a copy of the query mutated at rate p
              (2:1:1) written over it at a uniform offset, so every target is
              exactly Lt bp.
  wfa_mixed   n pairs; length log-uniform in [lo, hi]; divergence uniform in
              [0, dmax] per pair (2:1:1); smax = min(0.2 * len, cap).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


@dataclass
class PairSet:
    a: np.ndarray          # packed bytes, first sequence of each pair (WFA: pattern; SW: text)
    a_off: np.ndarray      # int64
    a_len: np.ndarray      # int32
    b: np.ndarray          # packed bytes, second sequence (WFA: text; SW: pattern)
    b_off: np.ndarray
    b_len: np.ndarray
    smax: Optional[np.ndarray] = None      # int32 per pair (WFA)
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.a_len.shape[0])

    def a_seq(self, i: int) -> bytes:
        o = int(self.a_off[i])
        return self.a[o: o + int(self.a_len[i])].tobytes()

    def b_seq(self, i: int) -> bytes:
        o = int(self.b_off[i])
        return self.b[o: o + int(self.b_len[i])].tobytes()

    def subset(self, idx) -> "PairSet":
        idx = np.asarray(idx)
        a, ao, al = _gather(self.a, self.a_off, self.a_len, idx)
        b, bo, bl = _gather(self.b, self.b_off, self.b_len, idx)
        sm = None if self.smax is None else self.smax[idx].copy()
        return PairSet(a, ao, al, b, bo, bl, sm, dict(self.meta))


def _gather(buf, off, ln, idx):
    lens = ln[idx].astype(np.int64)
    new_off = np.zeros(len(idx), dtype=np.int64)
    if len(idx) > 1:
        new_off[1:] = np.cumsum(lens)[:-1]
    total = int(lens.sum())
    src = np.repeat(off[idx] - new_off, lens) + np.arange(total, dtype=np.int64)
    return buf[src].copy(), new_off, ln[idx].astype(np.int32)


def _offsets(lens: np.ndarray) -> np.ndarray:
    off = np.zeros(len(lens), dtype=np.int64)
    if len(lens) > 1:
        off[1:] = np.cumsum(lens.astype(np.int64))[:-1]
    return off


def random_bases(rng: np.random.Generator, n: int) -> np.ndarray:
    return ACGT[rng.integers(0, 4, size=n, dtype=np.uint8)]


_CODE = np.zeros(256, dtype=np.uint8)
_CODE[ACGT] = np.arange(4, dtype=np.uint8)


def mutate(rng: np.random.Generator, seq: np.ndarray, lens: np.ndarray, rate: np.ndarray,
           chunk: int = 1 << 24):
    """Apply per-base edits (sub:ins:del = 2:1:1 at total probability rate[pair])
    to every sequence of a packed array.  Returns (bytes, offsets, lengths).
    Works through the pairs in groups of about `chunk` bases."""
    lens = lens.astype(np.int64)
    off = _offsets(lens)
    rate = np.asarray(rate, dtype=np.float32)
    n = len(lens)
    pieces, out_len = [], np.zeros(n, dtype=np.int64)
    ends = np.cumsum(lens)
    i = 0
    while i < n:
        j = int(np.searchsorted(ends, (ends[i - 1] if i else 0) + chunk, side="right"))
        j = max(j, i + 1)
        lo, hi = int(off[i]), int(off[j - 1] + lens[j - 1])
        total = hi - lo
        pair = np.repeat(np.arange(j - i), lens[i:j])
        r = rng.random(total, dtype=np.float32)
        p = rate[i:j][pair]
        sub = r < p * 0.5
        ins = (r >= p * 0.5) & (r < p * 0.75)
        dele = (r >= p * 0.75) & (r < p)
        base = seq[lo:hi].copy()
        # substitution: one of the three other bases
        shift = rng.integers(1, 4, size=total, dtype=np.uint8)
        base[sub] = ACGT[(_CODE[base[sub]] + shift[sub]) & 3]
        count = 1 + ins.astype(np.int64) - dele.astype(np.int64)
        out_len[i:j] = np.bincount(pair, weights=count, minlength=j - i).astype(np.int64)
        src = np.repeat(np.arange(total), count)
        out = base[src]
        # the second copy of an inserted base becomes a uniform random base
        dup = np.zeros(len(src), dtype=bool)
        dup[1:] = src[1:] == src[:-1]
        out[dup] = random_bases(rng, int(dup.sum()))
        pieces.append(out)
        i = j
    out = np.concatenate(pieces) if pieces else np.zeros(0, dtype=np.uint8)
    return out, _offsets(out_len), out_len.astype(np.int32)


def wfa_plain(n: int = 10_000, length: int = 150, rate: float = 0.02, smax: int = 64,
              seed: int = 20261202) -> PairSet:
    rng = np.random.default_rng(seed)
    pl = np.full(n, length, dtype=np.int32)
    pat = random_bases(rng, n * length)
    txt, toff, tl = mutate(rng, pat, pl, np.full(n, rate))
    return PairSet(pat, _offsets(pl), pl, txt, toff, tl, np.full(n, smax, dtype=np.int32),
                   meta={"workload": "wfa_plain", "n": n, "length": length, "rate": rate,
                         "smax": smax, "seed": seed})


def wfa_routing(n: int = 200_000, lo: int = 100, hi: int = 10_000, smax_frac: float = 0.2,
                smax_cap: int = 2048, seed: int = 20261205) -> PairSet:
    """Pairs for the learned router: lengths log-uniform lo..hi; 80% of pairs
    with divergence U[0, 0.03], 15% U[0.03, 0.15] (edits 2:1:1), 5% unrelated
    (the text is an independent random sequence whose length is drawn from the
    same distribution); smax = min(smax_frac * pattern length, smax_cap)."""
    rng = np.random.default_rng(seed)
    pl = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int32)
    kind = rng.choice(3, size=n, p=[0.80, 0.15, 0.05])
    rate = np.where(kind == 0, rng.uniform(0.0, 0.03, size=n), rng.uniform(0.03, 0.15, size=n))
    pat = random_bases(rng, int(pl.astype(np.int64).sum()))
    txt, toff, tl = mutate(rng, pat, pl, rate)
    unrel = np.flatnonzero(kind == 2)
    if len(unrel):
        ul = np.exp(rng.uniform(np.log(lo), np.log(hi), size=len(unrel))).astype(np.int32)
        ub = random_bases(rng, int(ul.astype(np.int64).sum()))
        # rebuild the text array with the unrelated texts swapped in
        lens = tl.astype(np.int64).copy()
        lens[unrel] = ul
        new_off = _offsets(lens)
        out = np.empty(int(lens.sum()), dtype=np.uint8)
        keep = np.ones(n, dtype=bool)
        keep[unrel] = False
        ki = np.flatnonzero(keep)
        kl = lens[ki]
        within = np.arange(int(kl.sum()), dtype=np.int64) - np.repeat(_offsets(kl), kl)
        out[np.repeat(new_off[ki], kl) + within] = txt[np.repeat(toff[ki], kl) + within]
        ull = ul.astype(np.int64)
        within = np.arange(int(ull.sum()), dtype=np.int64) - np.repeat(_offsets(ull), ull)
        out[np.repeat(new_off[unrel], ull) + within] = ub
        txt, toff, tl = out, new_off, lens.astype(np.int32)
        rate = rate.copy()
        rate[unrel] = np.nan
    smax = np.minimum((smax_frac * pl).astype(np.int32), smax_cap).astype(np.int32)
    return PairSet(pat, _offsets(pl), pl, txt, toff, tl, smax,
                   meta={"workload": "wfa_routing", "n": n, "lo": lo, "hi": hi, "seed": seed,
                         "kind": kind, "rate": rate})


# Marginal residue frequencies of the BLOSUM62 target distribution, in the
# matrix's ARNDCQEGHILKMFPSTWYV order.
PROTEIN = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
BLOSUM62_FREQ = np.array([0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
                          0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073])
BLOSUM62_FREQ = BLOSUM62_FREQ / BLOSUM62_FREQ.sum()


def protein_pairs(n: int = 2000, qlo: int = 100, qhi: int = 1000, tlo: int = 100, thi: int = 2000,
                  seed: int = 20261207) -> PairSet:
    """Protein SW pairs: a = target (text, rows) of length U[tlo, thi], b = query
    (pattern, columns) of length U[qlo, qhi]; residues i.i.d. at the BLOSUM62
    background frequencies."""
    rng = np.random.default_rng(seed)
    tl = rng.integers(tlo, thi + 1, size=n).astype(np.int32)
    ql = rng.integers(qlo, qhi + 1, size=n).astype(np.int32)
    t = PROTEIN[rng.choice(20, size=int(tl.astype(np.int64).sum()), p=BLOSUM62_FREQ)]
    q = PROTEIN[rng.choice(20, size=int(ql.astype(np.int64).sum()), p=BLOSUM62_FREQ)]
    return PairSet(t, _offsets(tl), tl, q, _offsets(ql), ql,
                   meta={"workload": "protein_pairs", "n": n, "seed": seed})


def sw_embedded(n: int = 50_000, qlen: int = 256, tlen: int = 1024, rate: float = 0.10,
                seed: int = 20261203) -> PairSet:
    rng = np.random.default_rng(seed)
    ql = np.full(n, qlen, dtype=np.int32)
    q = random_bases(rng, n * qlen)
    mut, moff, ml = mutate(rng, q, ql, np.full(n, rate))
    ml = np.minimum(ml, tlen)
    t = random_bases(rng, n * tlen).reshape(n, tlen)
    off = (rng.random(n) * (tlen - ml + 1)).astype(np.int64)
    rows = np.repeat(np.arange(n), ml)
    within = np.arange(int(ml.sum())) - np.repeat(_offsets(ml), ml)
    t[rows, np.repeat(off, ml) + within] = mut[np.repeat(moff, ml) + within]
    tl = np.full(n, tlen, dtype=np.int32)
    return PairSet(t.reshape(-1), _offsets(tl), tl, q, _offsets(ql), ql,
                   meta={"workload": "sw_embedded", "n": n, "qlen": qlen, "tlen": tlen,
                         "rate": rate, "seed": seed, "offset": off})


def wfa_mixed(n: int = 200_000, lo: int = 100, hi: int = 10_000, dmax: float = 0.15,
              smax_frac: float = 0.2, smax_cap: int = 2048, seed: int = 20261204,
              div: Optional[np.ndarray] = None) -> PairSet:
    rng = np.random.default_rng(seed)
    pl = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int32)
    rate = rng.uniform(0.0, dmax, size=n) if div is None else div
    pat = random_bases(rng, int(pl.astype(np.int64).sum()))
    txt, toff, tl = mutate(rng, pat, pl, rate)
    smax = np.minimum((smax_frac * pl).astype(np.int32), smax_cap).astype(np.int32)
    return PairSet(pat, _offsets(pl), pl, txt, toff, tl, smax,
                   meta={"workload": "wfa_mixed", "n": n, "lo": lo, "hi": hi, "dmax": dmax,
                         "seed": seed, "rate": rate})


# ---------------------------------------------------------------- serving
def serving_request(rng: np.random.Generator, api_mod, max_pairs: int = 512, wfa_frac: float = 0.6,
                    wfa_len=(100, 5000), wfa_div: float = 0.10, sw_q=(100, 1000), sw_t=(100, 2000),
                    cigar_frac: float = 0.7):
    """One request of the serving mix: (kind, items, with_cigar).  WFA: DNA
    pairs, lengths log-uniform in wfa_len, divergence U[0, wfa_div] (2:1:1),
    smax = min(0.2 len, 2048).  SW: protein pairs (query / target lengths
    uniform in sw_q / sw_t, BLOSUM62 background frequencies)."""
    k = int(rng.integers(1, max_pairs + 1))
    cig = bool(rng.random() < cigar_frac)
    if rng.random() < wfa_frac:
        pl = np.exp(rng.uniform(np.log(wfa_len[0]), np.log(wfa_len[1]), size=k)).astype(np.int32)
        pat = random_bases(rng, int(pl.astype(np.int64).sum()))
        txt, toff, tl = mutate(rng, pat, pl, rng.uniform(0.0, wfa_div, size=k))
        po = _offsets(pl)
        items = []
        for i in range(k):
            p = pat[po[i]: po[i] + pl[i]].tobytes()
            t = txt[toff[i]: toff[i] + tl[i]].tobytes()
            items.append(api_mod.AlignRequest(pattern=p, pattern_len=len(p), text=t, text_len=len(t),
                                              smax=int(min(int(0.2 * len(p)), 2048)), with_cigar=cig))
        return "wfa", items, cig
    ps = protein_pairs(n=k, qlo=sw_q[0], qhi=sw_q[1], tlo=sw_t[0], thi=sw_t[1],
                       seed=int(rng.integers(1 << 62)))
    items = [api_mod.SWRequest(text=ps.a_seq(i), text_len=int(ps.a_len[i]),
                               pattern=ps.b_seq(i), pattern_len=int(ps.b_len[i]), with_cigar=cig)
             for i in range(k)]
    return "sw", items, cig


def fasta_genome(total: int = 512 << 20, seed: int = 0, width: int = 60, rec_lo: int = 10_000,
                 rec_hi: int = 2_000_000) -> np.ndarray:
    """A synthetic multi-record FASTA file of about ``total`` bytes (uint8):
    records of uniform length in [rec_lo, rec_hi), ``width``-byte lines, uniform
    ACGT with the variations real assemblies carry: every third record 2 % N,
    every fifth a soft-masked (lowercase) first quarter, every seventh CRLF
    line ends, headers with a description."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    parts, size, k = [], 0, 0
    while size < total:
        ln = int(rng.integers(rec_lo, rec_hi))
        s = acgt[rng.integers(0, 4, ln, dtype=np.uint8)]
        if k % 3 == 1:
            s[rng.integers(0, ln, ln // 50)] = ord("N")
        if k % 5 == 2:
            s[: ln // 4] += 32
        eol = np.frombuffer(b"\r\n" if k % 7 == 3 else b"\n", np.uint8)
        nl = (ln + width - 1) // width
        pad = nl * width - ln
        body = np.concatenate([s, np.zeros(pad, np.uint8)]).reshape(nl, width)
        body = np.concatenate([body, np.tile(eol, (nl, 1))], axis=1).ravel()
        if pad:
            body = np.delete(body, np.s_[len(body) - len(eol) - pad:len(body) - len(eol)])
        hdr = np.frombuffer(f">chr{k} synthetic record {k} length={ln}\n".encode(), np.uint8)
        parts += [hdr, body]
        size += len(hdr) + len(body)
        k += 1
    return np.concatenate(parts)


_STOPS_ANY = {b"TAA", b"TAG", b"TGA", b"AGA", b"AGG"}


def gene_family(n: int = 5000, codons: int = 380, seed: int = 0):
    """A synthetic protein-coding gene family in FASTA form: (records file
    bytes, bait file bytes).  The bait is ATG + random sense codons (no stop
    under code 1 or 2) + TAA, like a mitochondrial CYTB (1,140 bp at 380
    codons).  The bait file holds a stop-riddled fragment, the bait itself and
    the bait behind one extra base (the same stop-free frame at offset 1, one
    base longer).  Records: 55 % the bait at 3 % substitutions, 10 % plus a 1-
    or 2-base indel (frameshift), 5 % plus an in-frame 3-base deletion, 10 %
    with a sense codon turned into TAG, 8 % fragments of 30-70 %, 5 % inside
    long random flanks, 3 % unrelated, 2 % soft-masked (lowercase), 2 % with N
    runs."""
    rng = np.random.default_rng(seed)
    bases = np.frombuffer(b"ACGT", np.uint8)
    sense = [bytes([a, b, c]) for a in b"ACGT" for b in b"ACGT" for c in b"ACGT"]
    sense = [s for s in sense if s not in _STOPS_ANY]
    bait = b"ATG" + b"".join(sense[k] for k in rng.integers(0, len(sense), codons - 2)) + b"TAA"

    def rand(k):
        return bases[rng.integers(0, 4, k)].tobytes()

    def subst(s, rate):
        a = np.frombuffer(s, np.uint8).copy()
        hit = rng.random(len(a)) < rate
        a[hit] = bases[rng.integers(0, 4, int(hit.sum()))]
        return a.tobytes()

    recs = []
    for k in range(n):
        u = rng.random()
        s = subst(bait, 0.03)
        if u < 0.55:
            pass
        elif u < 0.65:
            p = int(rng.integers(30, len(s) - 30))
            s = s[:p] + rand(int(rng.integers(1, 3))) + s[p:] if rng.random() < 0.5 else \
                s[:p] + s[p + int(rng.integers(1, 3)):]
        elif u < 0.70:
            p = 3 * int(rng.integers(10, codons - 10))
            s = s[:p] + s[p + 3:]
        elif u < 0.80:
            p = 3 * int(rng.integers(5, codons - 5))
            s = s[:p] + b"TAG" + s[p + 3:]
        elif u < 0.88:
            ln = int(len(s) * rng.uniform(0.3, 0.7))
            p = int(rng.integers(0, len(s) - ln))
            s = s[p:p + ln]
        elif u < 0.93:
            s = rand(int(rng.integers(800, 1600))) + s + rand(int(rng.integers(800, 1600)))
        elif u < 0.96:
            s = rand(len(bait))
        elif u < 0.98:
            s = s.lower()
        else:
            p = int(rng.integers(0, len(s) - 40))
            s = s[:p] + b"N" * int(rng.integers(5, 40)) + s[p + 40:]
        recs.append(b">acc%05d synthetic gene record\n" % k + b"\n".join(s[i:i + 70] for i in range(0, len(s), 70)))
    frag = b"TAAA" * 60 + rand(200)          # period 4: a TAA in every frame
    ref = (b">frag stop-riddled fragment\n" + frag + b"\n>bait0 clean\n" + bait + b"\n>bait1 shifted\nC" +
           bait + b"\n")
    return b"\n".join(recs) + b"\n", ref


# ---------------------------------------------------------------- MSA
def aligned_family(rng: np.random.Generator, L: int, rows: int, sub: float = 0.06, gap: float = 0.03,
                   ins: float = 0.02) -> List[str]:
    """`rows` aligned rows of one family: an ancestor of L bases, per row and
    column a substitution (rate `sub`) or a gap ('-', rate `gap`), and
    insertion columns (rate `ins` per column) where one row holds a base and
    the others a gap -- the shape of a progressive merge's profile."""
    anc = ACGT[rng.integers(0, 4, L, dtype=np.uint8)]
    cols = rng.random(L) < ins
    width = L + int(cols.sum())
    out = np.empty((rows, width), dtype=np.uint8)
    keep = np.ones(width, dtype=bool)
    pos = np.arange(L) + np.cumsum(cols)           # ancestor column positions in the aligned rows
    keep[:] = False
    keep[pos] = True
    for r in range(rows):
        row = np.full(width, ord("-"), dtype=np.uint8)
        b = anc.copy()
        m = rng.random(L)
        s = m < sub
        b[s] = ACGT[(_CODE[b[s]] + rng.integers(1, 4, int(s.sum()), dtype=np.uint8)) & 3]
        b[(m >= sub) & (m < sub + gap)] = ord("-")
        row[pos] = b
        ins_cols = np.flatnonzero(~keep)
        owner = rng.integers(0, rows, len(ins_cols))
        mine = ins_cols[owner == r]
        row[mine] = ACGT[rng.integers(0, 4, len(mine), dtype=np.uint8)]
        out[r] = row
    # drop columns that are gaps in every row
    nz = (out != ord("-")).any(axis=0)
    return [bytes(x).decode() for x in out[:, nz]]


def msa_profile_pairs(n: int = 4096, lo: int = 600, hi: int = 1400, rows=(1, 4), seed: int = 20261208):
    """Profile pairs of a progressive-MSA level: pair q aligns two profiles
    of one family (1..4 rows each, lengths uniform lo..hi), returned as row
    lists [(rows_a, rows_b)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        ra, rb = int(rng.integers(rows[0], rows[1] + 1)), int(rng.integers(rows[0], rows[1] + 1))
        fam = aligned_family(rng, L, ra + rb)
        a = [r for r in fam[:ra]]
        b = [r for r in fam[ra:]]
        # each side keeps only its own non-gap columns (a profile of its rows)
        def squeeze(rs):
            arr = np.frombuffer("".join(rs).encode(), np.uint8).reshape(len(rs), -1)
            k = (arr != ord("-")).any(axis=0)
            return [bytes(x).decode() for x in arr[:, k]]
        out.append((squeeze(a), squeeze(b)))
    return out


def sequence_family_set(n: int = 1000, L: int = 1200, families: int = 40, rate: float = 0.12,
                        seed: int = 20261209) -> List[str]:
    """n DNA sequences from `families` ancestors (substitutions and indels
    at up to `rate`), the input of a guide tree."""
    rng = np.random.default_rng(seed)
    ancs = [ACGT[rng.integers(0, 4, L, dtype=np.uint8)] for _ in range(families)]
    out = []
    for q in range(n):
        a = ancs[q % families]
        r = rng.uniform(0.0, rate)
        m = rng.random(L)
        b = a.copy()
        s = m < r * 0.6
        b[s] = ACGT[(_CODE[b[s]] + rng.integers(1, 4, int(s.sum()), dtype=np.uint8)) & 3]
        keep = ~((m >= r * 0.6) & (m < r * 0.8))
        out.append(bytes(b[keep]).decode())
    return out
