"""Real sequence on the GPU path: the reference's real-sequence cases
(tests/real/test_real_sequences.cpp:130-272) and its batch-smax case on
mtDNA (tests/api/test_batch_semantics.cpp:60-153), on the reference's own
data files (NC_012920.1 human and NC_001643.1 chimp mtDNA, copied under
tests/golden/).  Tandem repeats, homopolymer runs, the tie-heavy D-loop, a
genuine N and real human/chimp divergence reach the WFA kernels here; every
score and CIGAR is compared with the oracle exactly, and the unresolved cases
are checked against the Levenshtein distance."""
import os

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mtdna(oracle):
    out = []
    for name in ("mtdna_human.fa", "mtdna_chimp.fa"):
        recs, err = oracle.fasta_parse(open(os.path.join(GOLDEN, name), "rb").read())
        assert err is None and len(recs) == 1
        out.append(recs[0][2].decode("latin-1"))
    hum, chi = out
    assert (len(hum), len(chi)) == (16569, 16554)
    return hum, chi


def _req(ga, p, t, smax, cig=True):
    return ga.AlignRequest(pattern=p, pattern_len=len(p), text=t, text_len=len(t), smax=smax,
                           with_cigar=cig)


def _compare(ga, oracle, pairs, smax, cig=True, ext=False):
    """One batch (align_batch, or align_batch_ext with per-pair smax) against
    the oracle; returns the number of resolved pairs."""
    sm = smax if isinstance(smax, list) else [smax] * len(pairs)
    reqs = [_req(ga, p, t, s, cig) for (p, t), s in zip(pairs, sm)]
    b = (ga.align_batch_ext if ext else ga.align_batch)(reqs)
    assert b.ok(), b.error
    resolved = 0
    for (p, t), s, r in zip(pairs, sm, b.results):
        score, cigar = oracle.wfa(p, t, s)
        d = oracle.levenshtein(p, t)
        assert r.score == score and r.resolved == (score >= 0)
        assert (score == d) if d <= s else (score == -1)
        if cig and r.resolved:
            assert r.cigar == cigar
            if cigar:
                assert r.rescore_ok and r.wellformed_ok
        else:
            assert r.cigar == ""
        resolved += r.resolved
    assert (b.resolved_count, b.unresolved_count) == (resolved, len(pairs) - resolved)
    return resolved


def test_real_fragments_identical_and_two_substitutions(gpu, oracle, mtdna):
    import genoaligner_amd as ga
    hum, _ = mtdna
    pairs = []
    for s in (1000, 3000, 4500, 6000, 9000, 12000, 15000):
        frag = hum[s: s + 400]
        mut = list(frag)
        mut[10] = "C" if mut[10] == "A" else "A"
        mut[200] = "T" if mut[200] == "G" else "G"
        pairs += [(frag, frag), (frag, "".join(mut))]
    assert _compare(ga, oracle, pairs, 64) == len(pairs)


def test_dloop_windows(gpu, oracle, mtdna):
    """The control region across the origin: self, a 3-base rotation and a
    substitution every 23 bases, at smax 64 with CIGAR and through the
    per-pair entry point at larger bounds."""
    import genoaligner_amd as ga
    hum, _ = mtdna
    dloop = hum[16000:16569] + hum[0:600]
    shifted = dloop[3:] + dloop[:3]
    mut = list(dloop)
    for i in range(0, len(mut) - 7, 23):
        mut[i] = "C" if mut[i] == "A" else "A"
    pairs = [(dloop, dloop), (dloop, shifted), (dloop, "".join(mut))]
    _compare(ga, oracle, pairs, 64)
    assert _compare(ga, oracle, pairs, [64, 200, 511], ext=True) == 3


def test_human_vs_chimp_colinear_windows(gpu, oracle, mtdna):
    """600 bp windows (distance ~300: unresolved at 64 and 200), the same
    windows resolved through the per-pair entry point at smax 400, and 100 bp
    windows that must resolve at 64."""
    import genoaligner_amd as ga
    hum, chi = mtdna
    wins = (500, 2000, 4000, 8000, 11000, 14000)
    long_pairs = [(hum[w: w + 600], chi[w: w + 600]) for w in wins]
    assert _compare(ga, oracle, long_pairs, 64) == 0
    assert _compare(ga, oracle, long_pairs, 200, cig=False) == 0
    assert _compare(ga, oracle, long_pairs, 400, ext=True) == len(wins)
    short = [(hum[w: w + 100], chi[w: w + 100]) for w in wins]
    assert _compare(ga, oracle, short, 64) == len(wins)


def test_whole_mtdna(gpu, oracle, mtdna):
    """16.5 kbp molecules: human vs itself and vs chimp (distance 2,502) at
    smax 511 score-only (the reference's case), and through the per-pair entry
    point with CIGAR at smax 2,048 (unresolved), 2,501 (one short), 2,502 (at
    the distance) and 4,000 (the extension's limit)."""
    import genoaligner_amd as ga
    hum, chi = mtdna
    assert _compare(ga, oracle, [(hum, hum), (hum, chi)], 511, cig=False) == 1
    pairs = [(hum, hum), (hum, chi), (hum, chi), (hum, chi), (hum, chi), (chi, hum)]
    assert _compare(ga, oracle, pairs, [2048, 2048, 2501, 2502, 4000, 4000], ext=True) == 4


def test_window_around_the_n(gpu, oracle, mtdna):
    """The one N of human mtDNA (position 3,106): a 200 bp window around it and
    its reverse, score-only at 200 as the reference runs it, and with CIGAR
    through the per-pair entry point (the N matches only N)."""
    import genoaligner_amd as ga
    hum, _ = mtdna
    npos = hum.find("N")
    assert npos == 3106
    p = hum[npos - 100: npos + 100]
    pairs = [(p, p), (p, p[::-1])]
    _compare(ga, oracle, pairs, 200, cig=False)
    assert _compare(ga, oracle, pairs, [200, 200], ext=True) == 2
    mut = p[:100] + "A" + p[101:]
    assert _compare(ga, oracle, [(p, mut), (mut, p)], 16) == 2


def test_batch_uses_the_minimum_smax_on_mtdna(gpu, oracle, mtdna):
    """tests/api/test_batch_semantics.cpp:60-153: two 300 bp windows with a
    substitution every 15 bases (distance 20 each); a batch of smax 200 and 8
    resolves neither (the minimum governs), the wide request alone resolves,
    and grouping both at 200 resolves both."""
    import genoaligner_amd as ga
    hum, _ = mtdna
    w1, w2 = hum[1000:1300], hum[5000:5300]

    def every15(w):
        m = list(w)
        for i in range(0, len(m) - 5, 15):
            m[i] = "C" if m[i] == "A" else "A"
        return "".join(m)

    m1, m2 = every15(w1), every15(w2)
    d1, d2 = oracle.levenshtein(w1, m1), oracle.levenshtein(w2, m2)
    assert 8 < d1 <= 200 and 8 < d2 <= 200
    b = ga.align_batch([_req(ga, w1, m1, 200), _req(ga, w2, m2, 8)])
    assert b.ok() and (b.resolved_count, b.unresolved_count) == (0, 2)
    assert [r.score for r in b.results] == [-1, -1]
    alone = ga.align_batch([_req(ga, w1, m1, 200, False)])
    assert alone.resolved_count == 1 and alone.results[0].score == d1
    both = ga.align_batch([_req(ga, w1, m1, 200, False), _req(ga, w2, m2, 200, False)])
    assert both.resolved_count == 2 and [r.score for r in both.results] == [d1, d2]
    # the per-pair entry point keeps each request's own bound
    ext = ga.align_batch_ext([_req(ga, w1, m1, 200), _req(ga, w2, m2, 8)])
    assert [r.score for r in ext.results] == [d1, -1]
    assert ext.results[0].cigar == oracle.wfa(w1, m1, 200)[1]
