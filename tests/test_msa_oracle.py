"""The MSA oracle (oracle/ga_msa_oracle.c: profile counts, the
profile-profile Gotoh, k-mer distances, neighbour joining) pinned to the
reference: the golden vectors written from its own code
(tests/golden/msa_pp.tsv, kmer_nj.tsv), the reference compiled for the CPU
where it is built, and closed forms (k-mer sets recounted in Python, an
additive tree that neighbour joining must recover)."""
import random
import struct

import numpy as np
import pytest

from conftest import read_tsv


def f32(h: str) -> float:
    return struct.unpack("<f", struct.pack("<I", int(h, 16)))[0]


def msa_params(O, tag):
    return {
        "default": O.MsaParams(),
        "legacy": O.MsaParams(psgp=False),
        "global": O.MsaParams(free_end_gaps=False),
        "protein": O.MsaParams.protein(),
        "protein_semi": O.MsaParams(gap_open=11.0, gap_extend=1.0, free_end_gaps=True, alpha=20,
                                    sub=O.MSA_BLOSUM62.copy()),
    }[tag]


def golden_msa():
    """(tag, rows A, rows B, score, span, cigar)"""
    return [(r[0], r[1].split(","), r[2].split(","), f32(r[3]), tuple(int(x) for x in r[4:8]), r[8])
            for r in read_tsv("msa_pp.tsv")]


def golden_kmer_nj():
    """(k, alpha, seqs, D float32, children [n-1, 2])"""
    out = []
    for r in read_tsv("kmer_nj.tsv"):
        seqs = ["" if s == "-" else s for s in r[2].split(",")]
        D = np.array([f32(h) for h in r[3].split(",")] if r[3] else [], dtype=np.float32)
        ch = np.array([[int(x) for x in p.split(":")] for p in r[4].split(",")] if r[4] else [],
                      dtype=np.int32).reshape(-1, 2)
        out.append((int(r[0]), int(r[1]), seqs, D, ch))
    return out


def test_profile_align_matches_golden(oracle):
    rows = golden_msa()
    assert len(rows) > 350
    for tag, A, B, score, span, cigar in rows:
        P = msa_params(oracle, tag)
        got = oracle.msa_align(oracle.msa_profile(A, P.alpha), oracle.msa_profile(B, P.alpha), P)
        assert got == (score, span, cigar), (tag, A, B)
        # the CIGAR consumes exactly the span of each profile
        assert got[2].count("M") + got[2].count("I") == span[2] - span[0]
        assert got[2].count("M") + got[2].count("D") == span[3] - span[1]


def test_profile_counts_match_reference(oracle):
    if oracle.ref_msa() is None:
        pytest.skip("reference MSA build not present")
    rng = random.Random(5)
    for alpha, letters in ((4, "ACGTRYMKSWHBVDNXacgtn-U.*"), (20, "ARNDCQEGHILKMFPSTWYVBZJUXOacdx-*.")):
        for _ in range(40):
            L, n = rng.randint(1, 50), rng.randint(1, 7)
            rows = ["".join(rng.choice(letters) for _ in range(L)) for _ in range(n)]
            a, b = oracle.msa_profile(rows, alpha), oracle.ref_msa_profile(rows, alpha)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_profile_align_matches_reference_random(oracle):
    """Beyond the fixtures: random profiles (up to 9 rows, gap and ambiguity
    letters, 1..300 columns) under every parameter set, against the
    reference's align_profiles and its kernel body."""
    if oracle.ref_msa() is None:
        pytest.skip("reference MSA build not present")
    rng = random.Random(77)
    for it in range(150):
        tag = rng.choice(["default", "legacy", "global", "protein", "protein_semi"])
        P = msa_params(oracle, tag)
        letters = ("ACGT" * 6 + "RYN-") if P.alpha == 4 else ("ARNDCQEGHILKMFPSTWYV" * 2 + "BZX-")
        def prof():
            L, n = rng.randint(1, 300 if it % 10 == 0 else 60), rng.randint(1, 9)
            return ["".join(rng.choice(letters) for _ in range(L)) for _ in range(n)]
        A, B = prof(), prof()
        want = oracle.ref_msa_align(A, B, P)
        assert oracle.msa_align(oracle.msa_profile(A, P.alpha), oracle.msa_profile(B, P.alpha), P) == want
        if it % 5 == 0:
            assert oracle.ref_msa_align(A, B, P, kernel=True) == want


def test_kmer_distances_and_nj_match_golden(oracle):
    cases = golden_kmer_nj()
    assert len(cases) > 20
    for k, alpha, seqs, D, ch in cases:
        got = oracle.kmer_distances(seqs, k, alpha)
        assert np.array_equal(got, D), (k, alpha)
        if len(seqs) >= 2:
            assert np.array_equal(oracle.nj_tree(got, len(seqs)), ch), (k, alpha)


def test_kmer_distances_closed_form(oracle):
    """1 - |K_i & K_j| / min(|K_i|, |K_j|) over k-mer sets recounted in
    Python (runs of unambiguous letters only)."""
    rng = random.Random(3)
    idx = {c: i for i, c in enumerate("ACGT")}
    idx.update({c.lower(): i for c, i in list(idx.items())})
    idx["U"] = idx["u"] = 3
    for _ in range(20):
        k = rng.randint(2, 7)
        seqs = ["".join(rng.choice("ACGTACGTNRa") for _ in range(rng.randint(0, 80))) for _ in range(6)]
        sets = []
        for s in seqs:
            ks, run = set(), []
            for c in s:
                if c not in idx:
                    run = []
                    continue
                run.append(idx[c])
                if len(run) >= k:
                    ks.add(tuple(run[-k:]))
            sets.append(ks)
        want = []
        for i in range(6):
            for j in range(i + 1, 6):
                mn = min(len(sets[i]), len(sets[j]))
                want.append(np.float32(1.0) - np.float32(len(sets[i] & sets[j])) / np.float32(mn) if mn else 1.0)
        assert np.array_equal(oracle.kmer_distances(seqs, k), np.array(want, dtype=np.float32))


def test_nj_recovers_an_additive_tree(oracle):
    """Distances read off a tree with positive branch lengths are additive,
    and neighbour joining recovers that tree's cherries: ((0,1),(2,3)) joined
    to 4 and ((5,6),7)."""
    # leaf -> path of (edge id) to the root; edge lengths
    edges = {"a": 1.0, "b": 2.0, "c": 1.5, "d": 1.0, "e": 3.0, "f": 0.5, "g": 1.0, "h": 2.5, "i": 1.2,
             "j": 0.7, "k": 0.9, "l": 1.1, "m": 0.6}
    paths = {0: "a e j", 1: "b e j", 2: "c f j", 3: "d f j", 4: "g k", 5: "h i l", 6: "m i l", 7: "i2 l"}
    edges["i2"] = 2.0
    n = 8
    def dist(x, y):
        px, py = paths[x].split(), paths[y].split()
        common = set(px) & set(py)
        return sum(edges[e] for e in px if e not in common) + sum(edges[e] for e in py if e not in common)
    D = np.array([dist(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.float32)
    ch = oracle.nj_tree(D, n)
    cherries = {tuple(sorted(p)) for p in ch.tolist() if p[0] < n and p[1] < n}
    assert {(0, 1), (2, 3), (5, 6)} <= cherries
    assert ch.shape == (n - 1, 2) and sorted(ch.ravel().tolist()) == list(range(2 * n - 2))



def _alignments(m, n, last=None):
    """Every column-op string over M / I (consumes a column of A) / D (of B)
    with no I next to a D, aligning m columns of A with n of B."""
    if m == 0 and n == 0:
        yield ""
        return
    for op, dm, dn in (("M", 1, 1), ("I", 1, 0), ("D", 0, 1)):
        if m < dm or n < dn or {op, last} == {"I", "D"}:
            continue
        for rest in _alignments(m - dm, n - dn, op):
            yield op + rest


class _ClosedEndScorer:
    """Scores an alignment of two profiles under the closed-end rules of
    msa_ref.cpp:1063-1164, written from the alignment's runs, not from the
    DP: the alignment ends in an M column and no gap run touches a run of the
    other kind; an inner run of gaps over columns c0..c1 costs open(c0) plus
    ext(c) for c0 < c <= c1 (position-specific costs of the columns it
    spans, msa.hpp:123-134), and a leading run of L columns costs open of its
    L-th column plus (L - 1) * gap_extend (the boundary row and column).
    Column scores and costs in float64."""

    def __init__(self, A, B, P):
        (ca, self.oa), (cb, self.ob) = A, B
        if P.sub is not None:
            S = np.array(P.sub, dtype=np.float64).reshape(P.alpha, P.alpha)
        else:
            S = np.array([[P.match if x == y else (P.ts if {x, y} in ({0, 2}, {1, 3}) else P.tv)
                           for y in range(4)] for x in range(4)])
        self.col = ca.astype(np.float64) @ S @ cb.astype(np.float64).T
        self.P = P

    def open_(self, occ):
        P = self.P
        if not P.psgp or occ >= 1.0:
            return P.gap_open * occ
        return max(P.psgp_scale * P.gap_open * occ, P.psgp_min_open * P.gap_open)

    def ext_(self, occ):
        P = self.P
        if not P.psgp or occ >= 1.0:
            return P.gap_extend
        return max(P.gap_extend * occ, P.psgp_min_ext * P.gap_extend)

    def score(self, al):
        i = j = k = 0
        s = 0.0
        while k < len(al):
            r = k
            while r < len(al) and al[r] == al[k]:
                r += 1
            L = r - k
            if al[k] == "M":
                s += sum(self.col[i + t, j + t] for t in range(L))
                i += L
                j += L
            else:
                occ, c0 = (self.oa, i) if al[k] == "I" else (self.ob, j)
                if k == 0:
                    s -= self.open_(float(occ[c0 + L - 1])) + (L - 1) * self.P.gap_extend
                else:
                    s -= self.open_(float(occ[c0])) + sum(self.ext_(float(occ[c0 + t])) for t in range(1, L))
                if al[k] == "I":
                    i += L
                else:
                    j += L
            k = r
        return s

    def best(self):
        return max(self.score(al) for al in _alignments(len(self.oa), len(self.ob)) if al.endswith("M"))


@pytest.mark.parametrize("tag", ["global", "protein"])
def test_profile_align_brute_force_closed_ends(oracle, tag):
    """Closed end gaps, tiny profiles (one of 1..3 columns, the other of
    1..6, so inner gap runs are common; 1..3 rows with gaps and ambiguity
    codes): the oracle's score is the best over every alignment, enumerated
    and scored from its runs, and its own CIGAR scores exactly that."""
    P = msa_params(oracle, tag)
    letters = ("ACGT" * 3 + "RN--") if P.alpha == 4 else ("ARNDCQEGHILKMFPSTWYV" + "BX---")
    rng = random.Random(91 if tag == "global" else 92)
    for it in range(150):
        def prof(hi):
            L, n = rng.randint(1, hi), rng.randint(1, 3)
            return ["".join(rng.choice(letters) for _ in range(L)) for _ in range(n)]
        rA, rB = prof(3), prof(6)
        if it % 2:
            rA, rB = rB, rA
        A, B = oracle.msa_profile(rA, P.alpha), oracle.msa_profile(rB, P.alpha)
        score, span, cigar = oracle.msa_align(A, B, P)
        sc = _ClosedEndScorer(A, B, P)
        want = sc.best()
        tol = 1e-4 * max(1.0, abs(want))
        assert abs(score - want) <= tol, (score, want)
        assert span == (0, 0, len(A[1]), len(B[1]))
        assert cigar.endswith("M") and "ID" not in cigar and "DI" not in cigar
        assert abs(sc.score(cigar) - want) <= tol, (cigar, want)
