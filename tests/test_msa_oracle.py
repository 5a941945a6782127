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
