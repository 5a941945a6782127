"""The oracle (oracle/ga_oracle.c) pinned against the reference's golden
vectors, its known answers, independent mathematics, and -- where the
reference shim library is built -- the reference's own code on fresh inputs.

Pins per oracle function:
  ga_oracle_wfa           golden wfa_trace.tsv (score + CIGAR from the reference
                          kernel); Levenshtein DP; reference KATs; validators
  ga_oracle_wfa_score     golden flat scores; equals ga_oracle_wfa's score
  ga_oracle_levenshtein   brute-force recursion on tiny inputs; hand KATs
  ga_oracle_sw_trace      golden sw_trace.tsv; the independent value walk;
                          textbook Gotoh with -inf boundaries; reference KATs
  ga_oracle_sw_score_lane golden sw_score_w64.tsv (reference kernel at wave64)
  ga_oracle_sw_value_walk equals sw_trace on golden + random inputs
  validators              rescore/wellformed of every oracle CIGAR, and planted
                          corruptions that must be caught
"""
import functools
import random

import numpy as np

import pytest

from conftest import golden_sw_score, golden_sw_trace, golden_wfa


# ---------------------------------------------------------------- golden vectors
def test_wfa_golden(oracle):
    rows = golden_wfa()
    assert len(rows) > 700
    for smax, p, t, score, cigar, flat in rows:
        sc, cg = oracle.wfa(p, t, smax)
        assert (sc, cg) == (score, cigar), (smax, p, t)
        assert oracle.wfa_score(p, t, smax) == flat == score


def test_sw_trace_golden(oracle):
    rows = golden_sw_trace()
    assert len(rows) > 300
    for sc, t, p, score, si, sj, ei, ej, cigar in rows:
        assert oracle.sw_trace(t, p, sc) == (score, si, sj, ei, ej, cigar), (sc, t, p)


def test_sw_score_lane_golden(oracle):
    rows = golden_sw_score()
    assert len(rows) >= 60
    for sc, t, p, score, ei, ej in rows:
        assert oracle.sw_score_lane(t, p, sc, wsize=64) == (score, ei, ej), (sc, t, p)


# ---------------------------------------------------------------- known answers
# tests/parity/traceback_cpu.cpp:344-359 and tests/parity/r1_check.cpp:117-131
WFA_KATS = [
    ("ACGT", "ACGT", 0), ("ACGT", "ACGA", 1), ("ACGTA", "ACGT", 1), ("ACGT", "ACGTA", 1),
    ("GG", "TTT", 3), ("C", "GT", 2), ("CT", "CCC", 2), ("CG", "GAC", 3), ("A", "AA", 1),
    ("", "", 0), ("ACGT", "", 4), ("", "ACGT", 4), ("AAAA", "AAAAA", 1),
    ("ACGTACGT", "TTTTTTTT", 6), ("A", "A", 0), ("A", "T", 1), ("AA", "A", 1),
]


@pytest.mark.parametrize("p,t,want", WFA_KATS)
def test_wfa_kats(oracle, p, t, want):
    sc, cg = oracle.wfa(p, t, 64)
    assert sc == want
    assert oracle.levenshtein(p, t) == want
    if p or t:
        assert oracle.wfa_validate(cg, sc, p, t) == (True, True)


def test_wfa_api_examples(oracle):
    # tests/api/test_api.cpp:196-218 and :296-330
    assert oracle.wfa("ACGT", "ACGT", 8) == (0, "MMMM")
    assert oracle.wfa("", "ACGT", 8) == (4, "IIII")
    for L in (8, 40, 120, 300, 55):
        t = "".join("ACGT"[i % 4] for i in range(L))
        p = list(t)
        p[L // 2] = "C" if p[L // 2] == "A" else "A"
        sc, cg = oracle.wfa("".join(p), t, 64)
        assert sc == 1 and len(cg) == L


def test_sw_kats(oracle):
    # tests/api/test_api.cpp:104-119, :402-477, :596-604
    assert oracle.sw_trace("TTTACGTGTT", "ACGTGT", (2, -3, 5, 2)) == (12, 3, 0, 8, 5, "MMMMMM")
    assert oracle.sw_trace("ACGTACGT", "ACGTACGT", (2, -3, 5, 2)) == (16, 0, 0, 7, 7, "M" * 8)
    t = "AACCGGTT" + "C" * 6 + "TTGGCCAA"
    r = oracle.sw_trace(t, "AACCGGTTTTGGCCAA", (2, -3, 5, 2))
    assert r[0] == 17 and r[5] == "MMMMMMMMIIIIIIMMMMMMMM"
    assert oracle.sw_trace("AAAAAAAA", "CCCCCCCC", (2, -3, 5, 2)) == (0, -1, -1, -1, -1, "")
    assert oracle.sw_score_lane("TTTACGTGTT", "ACGTGT", (2, -3, 5, 2)) == (12, 8, 5)
    assert oracle.sw_score_lane("ACGT", "ACGT", (2, -3, 1, 2))[0] == 8


# tests/sw/test_sw_reference.cpp:147-215 under {1,-1,2,1}: the reference's
# hand-derived expectations
_CORE = "ACGTACGTACGT"
SW_REF_KATS = [
    ("T" * 10 + _CORE + "G" * 10, "C" * 10 + _CORE + "A" * 10, 12),
    ("ACGT", "ACGT", 4), ("A", "A", 1), ("A" * 20, "A" * 20, 20),
    ("AAAAAAAA", "CCCCCCCC", 0), ("ACGTACGT", "TGCATGCA", 1), ("AAAAAA", "CCCCCC", 0),
    ("A", "C", 0),
    ("ACGTACGT" + "T" * 20, "ACGTACGT" + "G" * 20, 8), ("T" * 20 + "ACGTACGT", "G" * 20 + "ACGTACGT", 8),
    ("ACGTTTACGT", "ACGTACGT", 5), ("ACACACAC", "ACAC", 4),
    ("", "ACGT", 0), ("ACGT", "", 0), ("", "", 0),
]


@pytest.mark.parametrize("t,p,want", SW_REF_KATS)
def test_sw_reference_kats(oracle, t, p, want):
    assert oracle.sw_trace(t, p, (1, -1, 2, 1))[0] == want


# ---------------------------------------------------------------- mathematics
def _lev_brute(a, b):
    @functools.lru_cache(None)
    def d(i, j):
        if i == 0:
            return j
        if j == 0:
            return i
        return min(d(i - 1, j) + 1, d(i, j - 1) + 1, d(i - 1, j - 1) + (a[i - 1] != b[j - 1]))
    return d(len(a), len(b))


def test_levenshtein_brute(oracle):
    rng = random.Random(5)
    for _ in range(300):
        a = "".join(rng.choice("AC") for _ in range(rng.randint(0, 6)))
        b = "".join(rng.choice("AC") for _ in range(rng.randint(0, 6)))
        assert oracle.levenshtein(a, b) == _lev_brute(a, b)


def _mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 2:
            out.append(rng.choice([x for x in "ACGT" if x != c]))
        elif r < rate * 3 / 4:
            out += [c, rng.choice("ACGT")]
        elif r >= rate:
            out.append(c)
    return "".join(out)


def test_wfa_score_is_bounded_edit_distance(oracle):
    """The WFA score is the Levenshtein distance when it is <= smax, and -1
    otherwise (degenerate pairs excepted); every CIGAR passes both checks."""
    rng = random.Random(11)
    for _ in range(1500):
        p = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 120)))
        t = _mutate(rng, p, rng.choice([0.0, 0.02, 0.1, 0.3, 0.8]))
        if not t:
            continue
        smax = rng.choice([0, 1, 3, 10, 40, 200])
        lev = oracle.levenshtein(p, t)
        sc, cg = oracle.wfa(p, t, smax)
        assert sc == (lev if lev <= smax else -1)
        if sc >= 0:
            assert oracle.wfa_validate(cg, sc, p, t) == (True, True)


def _gotoh_textbook(t, p, sc):
    """Textbook local Gotoh with -inf gap boundaries: the same optimal score
    as the reference's zero-boundary convention."""
    m, n = len(t), len(p)
    NEG = -10 ** 9
    match, mism, go, ge = sc
    H = [[0] * (n + 1) for _ in range(m + 1)]
    E = [[NEG] * (n + 1) for _ in range(m + 1)]
    F = [[NEG] * (n + 1) for _ in range(m + 1)]
    best = 0
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            E[i][j] = max(H[i][j - 1] - go, E[i][j - 1] - ge)
            F[i][j] = max(H[i - 1][j] - go, F[i - 1][j] - ge)
            H[i][j] = max(0, H[i - 1][j - 1] + (match if t[i - 1] == p[j - 1] else mism),
                          E[i][j], F[i][j])
            best = max(best, H[i][j])
    return best


def test_sw_scores_textbook_and_value_walk(oracle):
    rng = random.Random(13)
    for _ in range(400):
        sc = rng.choice([(2, -3, 5, 2), (1, -1, 2, 1), (2, -3, 3, 1), (3, -2, 4, 4)])
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        p = _mutate(rng, t[rng.randint(0, len(t)):], 0.2) if rng.random() < 0.6 else \
            "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        tr = oracle.sw_trace(t, p, sc)
        assert tr[0] == _gotoh_textbook(t, p, sc)
        assert oracle.sw_value_walk(t, p, sc) == tr
        assert oracle.sw_score_lane(t, p, sc)[0] == tr[0]
        if tr[0] > 0:
            assert oracle.sw_validate(tr[5], t, tr[1], tr[3], p, tr[2], tr[4], sc, tr[0]) == (True, True)


def test_sw_lane_rule_differs_from_row_major_on_ties(oracle):
    """Planted-difference check: on low-complexity inputs the wave64 lane rule
    and the row-major rule pick different end cells, so a path that swapped
    them would fail the golden comparison."""
    differs = 0
    for sc, t, p, score, ei, ej in golden_sw_score():
        tr = oracle.sw_trace(t, p, sc) if sc[3] <= sc[2] else None
        if tr and tr[0] > 0 and (tr[3], tr[4]) != (ei, ej):
            differs += 1
    assert differs > 0


# ---------------------------------------------------------------- validators
def test_validators_catch_corruption(oracle):
    sc, cg = oracle.wfa("ACGTACGTAC", "ACGTTCGTAC", 8)
    assert oracle.wfa_validate(cg, sc, "ACGTACGTAC", "ACGTTCGTAC") == (True, True)
    bad = cg.replace("X", "M", 1)
    assert oracle.wfa_validate(bad, sc, "ACGTACGTAC", "ACGTTCGTAC")[1] is False
    assert oracle.wfa_validate(cg + "I", sc, "ACGTACGTAC", "ACGTTCGTAC") == (False, False)
    r = oracle.sw_trace("TTTACGTGTT", "ACGTGT", (2, -3, 5, 2))
    assert oracle.sw_validate("MMMMMX", "TTTACGTGTT", 3, 8, "ACGTGT", 0, 5, (2, -3, 5, 2), 12) == (False, False)
    assert oracle.sw_validate(r[5], "TTTACGTGTT", 3, 8, "ACGTGT", 0, 5, (2, -3, 5, 2), 11) == (False, True)


# ---------------------------------------------------------------- batch rules
def test_batch_rules(oracle):
    from types import SimpleNamespace as R
    reqs = [R(pattern="ACGT", pattern_len=4, text="ACGA", text_len=4, smax=8, with_cigar=False),
            R(pattern="AAAA", pattern_len=4, text="TTTT", text_len=4, smax=2, with_cigar=True)]
    b = oracle.align_batch(reqs)
    assert b.status == "ok"
    # min smax = 2: the second pair (distance 4) is unresolved; any_cigar gives
    # the score-only request a CIGAR too
    assert (b.results[0].score, b.results[0].cigar) == (1, "MMMX")
    assert b.results[1].score == -1 and b.resolved_count == 1
    reqs[0].smax = 512
    assert oracle.align_batch(reqs).status == "invalid_argument"
    both_empty = oracle.align_batch([R(pattern="", pattern_len=0, text="", text_len=0, smax=4,
                                       with_cigar=True)]).results[0]
    assert (both_empty.score, both_empty.cigar, both_empty.rescore_ok) == (0, "", False)


# ---------------------------------------------------------------- the reference itself
needs_ref = pytest.mark.skipif("not __import__('oracle.oracle', fromlist=['x']).ref_available()",
                               reason="reference shim library not built")


@needs_ref
def test_oracle_equals_reference_random(oracle):
    rng = random.Random(99)
    for _ in range(800):
        p = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 80)))
        t = _mutate(rng, p.replace("N", "A"), rng.choice([0.0, 0.05, 0.3]))
        s = rng.choice([0, 2, 9, 33, 100])
        assert oracle.wfa(p, t, s) == oracle.ref_wfa(p, t, s)
    for _ in range(400):
        sc = rng.choice([(2, -3, 5, 2), (1, -1, 2, 1), (2, -3, 3, 1)])
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 60)))
        p = _mutate(rng, t[rng.randint(0, len(t)):], 0.15)
        assert oracle.sw_trace(t, p, sc) == oracle.ref_sw_trace(t, p, sc)


def test_matrix_table_reduces_to_match_mismatch(oracle):
    """A 4-letter ACGT matrix with 2 / -3 (unknown -3) gives the oracle exactly
    the match/mismatch table on every byte pair that matters: both sequences
    over ACGT plus other bytes (unequal bytes score the mismatch either way,
    equal non-ACGT bytes are the one difference and are left out)."""
    import random
    mm = oracle.mm_table(2, -3).reshape(256, 256)
    mt = oracle.matrix_table("ACGT", [[2 if a == b else -3 for b in range(4)] for a in range(4)], -3).reshape(256, 256)
    idx = [ord(c) for c in "ACGT"]
    assert (mm[np.ix_(idx, idx)] == mt[np.ix_(idx, idx)]).all()
    rng = random.Random(4)
    for _ in range(40):
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 60)))
        p = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 40)))
        assert oracle.sw_trace(t, p, (2, -3, 5, 2)) == oracle.sw_trace(t, p, mt, 5, 2)
