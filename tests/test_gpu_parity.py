"""GPU parity: the HIP path through the C ABI against the oracle, the golden
vectors written from the reference's own code, and the reference's result
rules.  Integer work, so every comparison is exact."""
import random

import numpy as np
import pytest

from conftest import golden_sw_score, golden_sw_trace, golden_wfa

pytestmark = pytest.mark.gpu


def _mutate(rng, s, rate, alpha="ACGT"):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 2:
            out.append(rng.choice([x for x in alpha if x != c] or [c]))
        elif r < rate * 3 / 4:
            out += [c, rng.choice(alpha)]
        elif r >= rate:
            out.append(c)
    return "".join(out)


def _areq(ga, p, t, smax, with_cigar=True):
    return ga.AlignRequest(pattern=p, pattern_len=len(p), text=t, text_len=len(t), smax=smax,
                           with_cigar=with_cigar)


def _sreq(ga, t, p, sc, with_cigar=True):
    return ga.SWRequest(text=t, text_len=len(t), pattern=p, pattern_len=len(p),
                        scoring=ga.SWScoring(*sc), with_cigar=with_cigar)


# ------------------------------------------------------------------ WFA
def test_wfa_golden_api(gpu):
    import genoaligner_amd as ga
    rows = golden_wfa()
    by_smax = {}
    for r in rows:
        by_smax.setdefault(r[0], []).append(r)
    checked = 0
    for smax, rs in sorted(by_smax.items()):
        b = ga.align_batch([_areq(ga, p, t, smax) for _, p, t, _, _, _ in rs])
        assert b.ok(), b.error
        for (s, p, t, score, cigar, _), res in zip(rs, b.results):
            assert (res.score, res.cigar) == (score, cigar), (s, p, t)
            if cigar:
                assert res.rescore_ok and res.wellformed_ok
            checked += 1
        # score-only batch: same scores, no CIGARs
        bs = ga.align_batch([_areq(ga, p, t, smax, False) for _, p, t, _, _, _ in rs])
        assert [r.score for r in bs.results] == [r[3] for r in rs]
        assert all(r.cigar == "" for r in bs.results)
    assert checked == len(rows)


def test_wfa_random_vs_oracle(gpu, oracle):
    import genoaligner_amd as ga
    rng = random.Random(2026)
    for smax in (0, 5, 64, 300, 511):
        reqs, pairs = [], []
        for _ in range(300):
            p = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 400)))
            t = _mutate(rng, p, rng.choice([0.0, 0.01, 0.05, 0.2, 0.6]))
            pairs.append((p, t))
            reqs.append(_areq(ga, p, t, smax))
        b = ga.align_batch(reqs)
        assert b.ok()
        for (p, t), res in zip(pairs, b.results):
            assert (res.score, res.cigar) == oracle.wfa(p, t, smax), (smax, p, t)


def test_wfa_batch_rules(gpu, oracle):
    import genoaligner_amd as ga
    # min smax and the mixed-batch CIGAR rule (src/api/api.cpp:271-274, :456-469)
    reqs = [_areq(ga, "ACGT", "ACGA", 8, False), _areq(ga, "AAAA", "TTTT", 2, True),
            _areq(ga, "", "", 5, False), _areq(ga, "ACGTNNacgt", "ACGTNNACGT", 9, False)]
    b = ga.align_batch(reqs)
    want = oracle.align_batch(reqs)
    assert b.ok()
    for got, exp in zip(b.results, want.results):
        assert (got.score, got.resolved, got.cigar, got.rescore_ok, got.wellformed_ok) == \
               (exp.score, exp.resolved, exp.cigar, exp.rescore_ok, exp.wellformed_ok)
    assert (b.resolved_count, b.unresolved_count) == (want.resolved_count, want.unresolved_count)
    # documented example (api.hpp:41-46)
    r = ga.align(_areq(ga, "ACGTACGT", "ACGTTCGT", 32))
    assert r.resolved and r.score == 1 and r.cigar == "MMMMXMMM"


def test_wfa_ext_per_pair_smax(gpu, oracle):
    import genoaligner_amd as ga
    rng = random.Random(7)
    reqs, exp = [], []
    for _ in range(200):
        L = rng.randint(50, 3000)
        p = "".join(rng.choice("ACGT") for _ in range(L))
        t = _mutate(rng, p, rng.uniform(0, 0.15))
        smax = min(int(0.2 * L), 2048)
        wc = rng.random() < 0.7
        reqs.append(_areq(ga, p, t, smax, wc))
        sc, cg = oracle.wfa(p, t, smax)
        exp.append((sc, cg if wc else ""))
    b = ga.align_batch_ext(reqs)
    assert b.ok()
    assert [(r.score, r.cigar) for r in b.results] == exp


def test_wfa_smax_at_the_distance(gpu, oracle):
    """Pairs whose bound is the exact distance, one more and one less (the
    last unresolved), with unequal lengths: the wavefront's budget pruning
    (diagonals that can no longer reach the end diagonal within smax are
    skipped) must leave every score and CIGAR as the full computation's."""
    import genoaligner_amd as ga
    rng = random.Random(11)
    reqs, exp = [], []
    for _ in range(120):
        L = rng.randint(40, 1500)
        p = "".join(rng.choice("ACGT") for _ in range(L))
        t = _mutate(rng, p, rng.uniform(0.02, 0.3))
        if rng.random() < 0.3:                    # shift the end diagonal
            t = t + "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 40)))
        d, _ = oracle.wfa(p, t, 4000)
        assert d >= 0
        for smax in (d - 1, d, d + 1):
            if smax < 1 or smax > 4000:
                continue
            reqs.append(_areq(ga, p, t, smax, True))
            exp.append(oracle.wfa(p, t, smax))
    b = ga.align_batch_ext(reqs)
    assert b.ok()
    got = [(r.score, r.cigar) for r in b.results]
    assert got == [(sc, cg if sc >= 0 else "") for sc, cg in exp]


def test_wfa_device_resident_plain(gpu, oracle):
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_plain(n=2000)
    dev = WfaDeviceBatch(ps, trace=True)
    dev.run()
    sc, cig = dev.results()
    for i in range(ps.n):
        p, t = ps.a_seq(i), ps.b_seq(i)
        osc, ocg = oracle.wfa(p, t, int(ps.smax[i]))
        assert (int(sc[i]), cig[i].decode()) == (osc, ocg), i


# ------------------------------------------------------------------ SW
def test_sw_trace_golden_api(gpu):
    import genoaligner_amd as ga
    rows = golden_sw_trace()
    by = {}
    for r in rows:
        by.setdefault(r[0], []).append(r)
    for sc, rs in by.items():
        b = ga.align_sw_batch([_sreq(ga, t, p, sc) for _, t, p, *_ in rs])
        assert b.ok(), b.error
        for (_, t, p, score, si, sj, ei, ej, cigar), res in zip(rs, b.results):
            assert (res.score, res.start_i, res.start_j, res.end_i, res.end_j, res.cigar) == \
                   (score, si, sj, ei, ej, cigar), (sc, t, p)
            assert res.resolved
            if score > 0:
                assert res.rescore_ok and res.wellformed_ok


def test_sw_score_only_golden_api(gpu):
    """Score-only batches report the end cell by the wave64 lane rule."""
    import genoaligner_amd as ga
    rows = golden_sw_score()
    by = {}
    for r in rows:
        by.setdefault(r[0], []).append(r)
    for sc, rs in by.items():
        b = ga.align_sw_batch([_sreq(ga, t, p, sc, False) for _, t, p, *_ in rs])
        assert b.ok(), b.error
        for (_, t, p, score, ei, ej), res in zip(rs, b.results):
            assert (res.score, res.end_i, res.end_j, res.start_i, res.start_j, res.cigar) == \
                   (score, ei, ej, -1, -1, ""), (sc, t, p)


def test_sw_random_vs_oracle(gpu, oracle):
    import genoaligner_amd as ga
    rng = random.Random(31)
    # (2, -5, 2, 1) and (3, -6, 4, 2): gap_open < -mismatch, where the packed
    # kernel's H carries the extra bias term
    for sc in ((2, -3, 5, 2), (2, -3, 3, 1), (1, -1, 2, 1), (2, -5, 2, 1), (3, -6, 4, 2)):
        reqs, pairs = [], []
        for _ in range(150):
            t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 700)))
            a = rng.randint(0, len(t))
            p = _mutate(rng, t[a: a + rng.randint(0, 300)], 0.1)
            if rng.random() < 0.2:
                p = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 200)))
            pairs.append((t, p))
            reqs.append(_sreq(ga, t, p, sc, rng.random() < 0.8))
        b = ga.align_sw_batch(reqs)
        want = oracle.align_sw_batch(reqs)
        assert b.ok()
        for got, exp, (t, p) in zip(b.results, want.results, pairs):
            assert (got.score, got.resolved, got.cigar, got.start_i, got.start_j, got.end_i,
                    got.end_j, got.rescore_ok, got.wellformed_ok, got.too_large) == \
                   (exp.score, exp.resolved, exp.cigar, exp.start_i, exp.start_j, exp.end_i,
                    exp.end_j, exp.rescore_ok, exp.wellformed_ok, exp.too_large), (sc, t, p)


def test_sw_long_gaps_all_trace_layouts(gpu, oracle):
    """Pairs with inserted and deleted blocks of 2..40 bases, under the three
    packed trace records, each with and without the gk bias term:
    match - mismatch <= 7 and gap_open - gap_extend <= 7 (H mod 8 and fe; the
    walk derives diagonal and gap steps from H mod 8), match - mismatch > 7
    with gap_open - gap_extend <= 3 (H mod 4 for the gap steps, H-source flags)
    and the rest (four flags)."""
    import genoaligner_amd as ga
    rng = random.Random(47)
    for sc in ((2, -3, 3, 1), (2, -3, 5, 2), (1, -2, 4, 1), (2, -3, 6, 1), (3, -4, 9, 2), (2, -5, 2, 1),
               (5, -4, 6, 4), (5, -4, 3, 1), (2, -8, 6, 1), (1, -7, 9, 1)):
        reqs, pairs = [], []
        for _ in range(120):
            t = "".join(rng.choice("ACGT") for _ in range(rng.randint(50, 900)))
            a = rng.randint(0, len(t) // 2)
            q = list(t[a: a + rng.randint(20, 300)])
            for _ in range(rng.randint(0, 4)):       # gap blocks in the query copy
                at = rng.randint(0, len(q))
                if rng.random() < 0.5:
                    q[at:at] = [rng.choice("ACGT") for _ in range(rng.randint(2, 40))]
                else:
                    del q[at: at + rng.randint(2, 40)]
            p = _mutate(rng, "".join(q), 0.05)
            pairs.append((t, p))
            reqs.append(_sreq(ga, t, p, sc, True))
        b = ga.align_sw_batch(reqs)
        want = oracle.align_sw_batch(reqs)
        assert b.ok()
        for got, exp, (t, p) in zip(b.results, want.results, pairs):
            assert (got.score, got.cigar, got.start_i, got.start_j, got.end_i, got.end_j) == \
                   (exp.score, exp.cigar, exp.start_i, exp.start_j, exp.end_i, exp.end_j), (sc, t, p)


def test_sw_run_walk_ties_and_borders(gpu, oracle):
    """The run walk over the H mod 8 record on the inputs where a traceback's
    tie rules decide the most: homopolymers, dinucleotide and tandem repeats
    with indels, alignments that start at row or column 1, pairs of 1..40
    bases, widths at the walk window's and the strip's seams (31..33, 63..65,
    127..129 columns) and texts up to 2,500 rows; under schemes at the
    record's limits (match - mismatch = 7, gap_open - gap_extend = 7), a
    positive mismatch, and gap_open == gap_extend."""
    import genoaligner_amd as ga
    rng = random.Random(53)
    schemes = ((2, -3, 3, 1), (1, -6, 8, 1), (4, -3, 9, 2), (3, 1, 2, 1), (5, -2, 5, 5), (1, -1, 1, 1))

    def rep(unit, n):
        return (unit * (n // len(unit) + 1))[:n]

    pairs = []
    for _ in range(12):                                  # homopolymers, repeats with indels
        for unit in ("A", "AC", "ACG", "AAT", "GATTACA"):
            t = rep(unit, rng.randint(20, 400))
            pairs.append((t, _mutate(rng, t[rng.randint(0, len(t) // 3):], 0.08)))
    for _ in range(40):                                  # alignments touching row / column 1
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(5, 300)))
        q = _mutate(rng, t[: rng.randint(1, len(t))], 0.1)
        pairs.append((t, q) if rng.random() < 0.5 else (q, t))
    for _ in range(80):                                  # tiny pairs
        pairs.append(("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 40))),
                      "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 40)))))
    for w in (31, 32, 33, 63, 64, 65, 127, 128, 129):    # window and strip seams
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(w, 3 * w)))
        a = rng.randint(0, len(t) - w)
        pairs.append((t, _mutate(rng, t[a: a + w], 0.06)[:w]))
    for _ in range(6):                                   # long texts, indel blocks
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(1500, 2500)))
        a = rng.randint(0, len(t) - 600)
        q = list(t[a: a + rng.randint(300, 600)])
        for _ in range(3):
            at = rng.randint(0, len(q))
            if rng.random() < 0.5:
                q[at:at] = [rng.choice("ACGT") for _ in range(rng.randint(2, 30))]
            else:
                del q[at: at + rng.randint(2, 30)]
        pairs.append((t, _mutate(rng, "".join(q), 0.05)))
    for sc in schemes:
        reqs = [_sreq(ga, t, p, sc, True) for t, p in pairs]
        b = ga.align_sw_batch(reqs)
        want = oracle.align_sw_batch(reqs)
        assert b.ok(), b.error
        for got, exp, (t, p) in zip(b.results, want.results, pairs):
            assert (got.score, got.cigar, got.start_i, got.start_j, got.end_i, got.end_j, got.rescore_ok,
                    got.wellformed_ok) == \
                   (exp.score, exp.cigar, exp.start_i, exp.start_j, exp.end_i, exp.end_j, exp.rescore_ok,
                    exp.wellformed_ok), (sc, t, p)


def test_sw_long_text_global_boundary(gpu, oracle):
    """Texts longer than the LDS boundary buffer take the HBM boundary path."""
    import genoaligner_amd as ga
    rng = random.Random(3)
    t = "".join(rng.choice("ACGT") for _ in range(9000))
    p = _mutate(rng, t[4000:4150], 0.1)
    reqs = [_sreq(ga, t, p, (2, -3, 3, 1)), _sreq(ga, t[:100], p, (2, -3, 3, 1))]
    b = ga.align_sw_batch(reqs)
    want = oracle.align_sw_batch(reqs)
    for got, exp in zip(b.results, want.results):
        assert (got.score, got.cigar, got.start_i, got.end_i, got.start_j, got.end_j) == \
               (exp.score, exp.cigar, exp.start_i, exp.end_i, exp.start_j, exp.end_j)


def test_sw_long_pattern_cigar(gpu, oracle):
    """Patterns longer than the packed kernel stages in LDS (8192 columns with
    CIGAR) take the int32 kernel; shorter ones in the same batch stay packed."""
    import genoaligner_amd as ga
    rng = random.Random(5)
    p = "".join(rng.choice("ACGT") for _ in range(9000))
    t = _mutate(rng, p[3000:3300], 0.1)
    reqs = [_sreq(ga, t, p, (2, -3, 3, 1)), _sreq(ga, t, p[2900:3400], (2, -3, 3, 1)),
            _sreq(ga, t[:80], p[:8192], (2, -3, 3, 1))]
    b = ga.align_sw_batch(reqs)
    want = oracle.align_sw_batch(reqs)
    assert b.ok(), b.error
    for got, exp in zip(b.results, want.results):
        assert (got.score, got.cigar, got.start_i, got.end_i, got.start_j, got.end_j) == \
               (exp.score, exp.cigar, exp.start_i, exp.end_i, exp.start_j, exp.end_j)


def test_sw_long_texts_packed_chunks(gpu, oracle):
    """Texts of 4097..8192 rows take the packed kernel (up to 64 KB of LDS
    boundary per wave, more with a long pattern's CIGAR staging), and a batch
    whose waves differ in LDS need runs as several chunks on two streams."""
    import genoaligner_amd as ga
    rng = random.Random(17)
    reqs = []
    # 8192 x 8000 with CIGAR: the walk's staging needs more than 64 KB of LDS
    t = "".join(rng.choice("ACGT") for _ in range(8192))
    reqs.append(_sreq(ga, t, _mutate(rng, t[100:8100], 0.05), (2, -3, 3, 1)))
    for lo, hi, k, pl in ((4097, 8192, 6, 1200), (2049, 4096, 40, 800), (100, 1000, 300, 300)):
        for _ in range(k):
            t = "".join(rng.choice("ACGT") for _ in range(rng.randint(lo, hi)))
            a = rng.randint(0, len(t) - 1)
            p = _mutate(rng, t[a: a + rng.randint(1, pl)], 0.1)
            if rng.random() < 0.15:
                p = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, pl)))
            reqs.append(_sreq(ga, t, p, (2, -3, 3, 1), rng.random() < 0.8))
    b = ga.align_sw_batch(reqs)
    want = oracle.align_sw_batch(reqs)
    assert b.ok(), b.error
    for k, (got, exp) in enumerate(zip(b.results, want.results)):
        assert (got.score, got.resolved, got.cigar, got.start_i, got.start_j, got.end_i,
                got.end_j, got.rescore_ok, got.wellformed_ok) == \
               (exp.score, exp.resolved, exp.cigar, exp.start_i, exp.start_j, exp.end_i,
                exp.end_j, exp.rescore_ok, exp.wellformed_ok), (k, len(reqs[k].text), len(reqs[k].pattern))


def test_sw_large_batch_piecewise_upload(gpu, oracle):
    """5,000 pairs through ga_align_sw_batch: the host packs and uploads the
    sequences in four pieces, the device checks the CIGARs and packs them for
    the copy back; every result, CIGAR and validator flag against the oracle,
    with score-only pairs, empty sequences and non-ACGT bytes mixed in."""
    import genoaligner_amd as ga
    rng = random.Random(2024)
    sc = (2, -3, 3, 1)
    reqs = []
    for k in range(5000):
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 420)))
        a = rng.randint(0, max(len(t) - 1, 0))
        p = _mutate(rng, t[a: a + rng.randint(0, 160)], 0.08)
        if rng.random() < 0.1:
            p = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 120)))
        if rng.random() < 0.03:
            t = t[:len(t) // 2] + "N" * 5 + t[len(t) // 2:]
        reqs.append(_sreq(ga, t, p, sc, rng.random() < 0.8))
    b = ga.align_sw_batch(reqs)
    want = oracle.align_sw_batch(reqs)
    assert b.ok(), b.error
    for k, (got, exp) in enumerate(zip(b.results, want.results)):
        assert (got.score, got.resolved, got.cigar, got.start_i, got.start_j, got.end_i, got.end_j,
                got.rescore_ok, got.wellformed_ok, got.too_large) == \
               (exp.score, exp.resolved, exp.cigar, exp.start_i, exp.start_j, exp.end_i, exp.end_j,
                exp.rescore_ok, exp.wellformed_ok, exp.too_large), k


def test_sw_too_large(gpu):
    import genoaligner_amd as ga
    big = "A" * 8192
    reqs = [_sreq(ga, big, big[:8193], (2, -3, 5, 2)), _sreq(ga, "TTTACGTGTT", "ACGTGT", (2, -3, 5, 2))]
    reqs[0] = _sreq(ga, big + "C", big, (2, -3, 5, 2))        # (8194)(8193) > 2^26 cells
    b = ga.align_sw_batch(reqs)
    assert b.ok()
    assert b.results[0].too_large and not b.results[0].resolved and b.results[0].score == -1
    r = b.results[1]
    assert (r.score, r.cigar, r.start_i, r.end_i, r.start_j, r.end_j) == (12, "MMMMMM", 3, 8, 0, 5)
    assert (b.resolved_count, b.unresolved_count) == (1, 1)


def test_sw_device_resident_embedded(gpu, oracle):
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import SwDeviceBatch
    ps = W.sw_embedded(n=256)
    dev = SwDeviceBatch(ps, scoring=(2, -3, 3, 1), trace=True)
    dev.run()
    out, cig = dev.results()
    for i in range(ps.n):
        t, p = ps.a_seq(i), ps.b_seq(i)
        r = oracle.sw_trace(t, p, (2, -3, 3, 1))
        assert tuple(int(x) for x in out[i]) == r[:5], i
        assert cig[i].decode() == r[5], i
    dev2 = SwDeviceBatch(ps, scoring=(2, -3, 3, 1), trace=False)
    dev2.run()
    out2, _ = dev2.results()
    for i in range(0, ps.n, 7):
        t, p = ps.a_seq(i), ps.b_seq(i)
        s, ei, ej = oracle.sw_score_lane(t, p, (2, -3, 3, 1))
        assert tuple(int(x) for x in out2[i]) == (s, -1, -1, ei, ej)


def test_sw_literal_bytes_pass(gpu, oracle):
    """Pairs holding bytes other than A/C/G/T (N runs, soft-masked lowercase,
    IUPAC codes, punctuation) mixed with plain pairs, so waves hold one of
    each: under schemes with mismatch < 0 the packed kernel's literal-byte
    pass runs them (every layout, with and without the gk term), under
    mismatch >= 0 the int32 kernel does; a pair of the literal-byte pass that
    outgrows 16 bits goes on to the int32 kernel.  Through the batch entry
    point (score-only and CIGAR batches) and the device entry point."""
    import genoaligner_amd as ga
    from genoaligner_amd.device import SwDeviceBatch
    from genoaligner_amd.workloads import PairSet
    import numpy as np
    rng = random.Random(59)
    pairs = []
    for k in range(160):
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(30, 900)))
        a = rng.randint(0, len(t) // 2)
        p = _mutate(rng, t[a: a + rng.randint(20, 250)], 0.08)
        if k % 2 == 0:                                  # every other pair: other bytes in it
            kind = k % 8
            tl, pl = list(t), list(p)
            for _ in range(rng.randint(1, 4)):
                if kind == 0:                           # an N run in the text
                    at = rng.randint(0, len(tl) - 1)
                    tl[at: at + rng.randint(1, 12)] = ["N"] * rng.randint(1, 12)
                elif kind == 2:                         # soft-masked run in both
                    at = rng.randint(0, len(tl) - 1)
                    tl[at: at + 20] = [c.lower() for c in tl[at: at + 20]]
                    bt = rng.randint(0, len(pl) - 1)
                    pl[bt: bt + 10] = [c.lower() for c in pl[bt: bt + 10]]
                elif kind == 4:                         # IUPAC codes in the pattern
                    pl[rng.randint(0, len(pl) - 1)] = rng.choice("RYKMSWN")
                else:                                   # other bytes on both sides
                    tl[rng.randint(0, len(tl) - 1)] = rng.choice("@#xX-")
                    pl[rng.randint(0, len(pl) - 1)] = rng.choice("@#xX-")
            t, p = "".join(tl), "".join(pl)
        pairs.append((t, p))
    big = "".join(rng.choice("ACGT") for _ in range(3500))
    pairs.append((big[:1700] + "N" + big[1701:], big[:1700] + "N" + big[1701:3000]))
    schemes = ((2, -3, 3, 1), (1, -1, 2, 1), (2, -5, 2, 1), (5, -4, 6, 4), (2, -8, 6, 1), (3, 1, 2, 1),
               (12, -3, 14, 1))
    for sc in schemes:
        for cig in (True, False):
            reqs = [_sreq(ga, t, p, sc, cig) for t, p in pairs]
            b = ga.align_sw_batch(reqs)
            want = oracle.align_sw_batch(reqs)
            assert b.ok(), b.error
            for got, exp, (t, p) in zip(b.results, want.results, pairs):
                assert (got.score, got.cigar, got.start_i, got.start_j, got.end_i, got.end_j, got.rescore_ok,
                        got.wellformed_ok) == \
                       (exp.score, exp.cigar, exp.start_i, exp.start_j, exp.end_i, exp.end_j, exp.rescore_ok,
                        exp.wellformed_ok), (sc, cig, t, p)
    # the device entry point (sequences in HBM) on the same pairs
    ta = [t.encode() for t, _ in pairs]
    pa = [p.encode() for _, p in pairs]
    tl = np.array([len(x) for x in ta], np.int32)
    pl = np.array([len(x) for x in pa], np.int32)
    to = np.zeros(len(ta), np.int64)
    po = np.zeros(len(pa), np.int64)
    to[1:] = np.cumsum(tl)[:-1]
    po[1:] = np.cumsum(pl)[:-1]
    ps = PairSet(np.frombuffer(b"".join(ta), np.uint8), to, tl, np.frombuffer(b"".join(pa), np.uint8), po, pl,
                 np.zeros(len(ta), np.int32), {})
    for sc in ((2, -3, 3, 1), (12, -3, 14, 1)):
        dev = SwDeviceBatch(ps, scoring=sc, trace=True)
        dev.run()
        out, cg = dev.results()
        for i, (t, p) in enumerate(pairs):
            r = oracle.sw_trace(t, p, sc)
            assert tuple(int(x) for x in out[i]) == r[:5], (sc, i)
            assert cg[i].decode() == r[5], (sc, i)


def test_sw_packed_overflow_rerun(gpu, oracle):
    """Scores beyond the packed kernel's 16-bit range are re-run in int32 and
    still match; neighbours in the same wave are unaffected."""
    import genoaligner_amd as ga
    rng = random.Random(17)
    big = "".join(rng.choice("ACGT") for _ in range(4000))
    sc = (10, -3, 12, 2)
    pairs = [(big, big), (big[:300], _mutate(rng, big[:300], 0.1)), (big[:3000], big[100:2900])]
    reqs = [_sreq(ga, t, p, sc) for t, p in pairs]
    b = ga.align_sw_batch(reqs)
    want = oracle.align_sw_batch(reqs)
    assert b.results[0].score == 40000
    for got, exp in zip(b.results, want.results):
        assert (got.score, got.cigar, got.start_i, got.end_i, got.start_j, got.end_j) == \
               (exp.score, exp.cigar, exp.start_i, exp.end_i, exp.start_j, exp.end_j)
    bs = ga.align_sw_batch([_sreq(ga, t, p, sc, False) for t, p in pairs])
    for got, (t, p) in zip(bs.results, pairs):
        assert (got.score, got.end_i, got.end_j) == oracle.sw_score_lane(t, p, sc)


def test_wfa_device_classes_want(gpu, oracle):
    """Mixed lengths and smax across size classes, 70% CIGAR / 30% score-only
    per pair, scheduled by class (longest first) and as one class."""
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_mixed(n=600, lo=20, hi=2500, seed=99)
    want = (np.random.default_rng(5).random(ps.n) < 0.7).astype(np.uint8)
    exp = [oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i])) for i in range(ps.n)]
    for classes in (True, False):
        dev = WfaDeviceBatch(ps, trace=True, want=want, classes=classes)
        dev.run()
        sc, cig = dev.results()
        for i in range(ps.n):
            e_sc, e_cg = exp[i]
            assert int(sc[i]) == e_sc, (classes, i)
            assert cig[i].decode() == (e_cg if want[i] else ""), (classes, i)
    dev = WfaDeviceBatch(ps, trace=False)
    dev.run()
    sc, _ = dev.results()
    assert [int(x) for x in sc] == [e[0] for e in exp]


def test_wfa_long_text_wide_stack(gpu, oracle):
    """Texts of 32767 bytes or more keep 32-bit wavefront stacks."""
    import genoaligner_amd as ga
    rng = random.Random(11)
    p = "".join(rng.choice("ACGT") for _ in range(33000))
    t = _mutate(rng, p, 0.002)
    reqs = [_areq(ga, p, t, 200), _areq(ga, p[:500], t[:505], 200), _areq(ga, t, p, 200, False)]
    b = ga.align_batch_ext(reqs)
    assert b.ok()
    for r, q in zip(b.results, reqs):
        sc, cg = oracle.wfa(q.pattern, q.text, 200)
        assert (r.score, r.cigar) == (sc, cg if q.with_cigar else "")


def test_wfa_non_acgt_bytes_across_classes(gpu, oracle):
    """Pairs holding N, lowercase or other bytes are compared byte for byte
    (N matches only N) in every size class; all-ACGT neighbours are unaffected."""
    import genoaligner_amd as ga
    rng = random.Random(23)
    reqs, exp = [], []
    for i in range(160):
        L = rng.choice([40, 300, 1200, 4000])
        p = "".join(rng.choice("ACGT") for _ in range(L))
        t = _mutate(rng, p, rng.uniform(0, 0.1))
        if i % 3 == 0:
            j = rng.randrange(len(t)) if t else 0
            t = t[:j] + rng.choice("NnacgtX") + t[j + 1:]
        if i % 5 == 0:
            j = rng.randrange(len(p))
            p = p[:j] + "N" + p[j + 1:]
        smax = min(int(0.2 * L) + 1, 2048)
        reqs.append(_areq(ga, p, t, smax, i % 4 != 0))
        sc, cg = oracle.wfa(p, t, smax)
        exp.append((sc, cg if i % 4 != 0 else ""))
    b = ga.align_batch_ext(reqs)
    assert b.ok()
    assert [(r.score, r.cigar) for r in b.results] == exp


# ------------------------------------------------------------------ protein SW
def _protein(rng, n, alphabet="ARNDCQEGHILKMFPSTWYV"):
    return "".join(rng.choice(alphabet) for _ in range(n))


def test_protein_sw_blosum62_vs_oracle(gpu, oracle):
    """BLOSUM62 with gap open 11 / extend 1: score, span and CIGAR (trace
    batches) and the lane-rule end (score-only batches) equal the oracle's
    table-driven Gotoh; residues outside the alphabet score -1."""
    import genoaligner_amd as ga
    m = ga.blosum62(11, 1)
    alpha = m.alphabet.decode()
    S = [[m.score[a][b] for b in range(20)] for a in range(20)]
    tab = oracle.matrix_table(alpha, S, -1)
    rng = random.Random(41)
    reqs, pairs = [], []
    for k in range(160):
        t = _protein(rng, rng.randint(0, 600))
        if rng.random() < 0.6 and t:
            a = rng.randint(0, len(t) - 1)
            p = _mutate(rng, t[a: a + rng.randint(1, 300)], 0.25, alpha)
        else:
            p = _protein(rng, rng.randint(0, 300))
        if k % 9 == 0 and p:
            p = p[:1] + "X" + p[2:]
        if k % 13 == 0 and t:
            t = "b" + t[1:]
        pairs.append((t, p))
        reqs.append(ga.SWRequest(text=t, text_len=len(t), pattern=p, pattern_len=len(p), with_cigar=True))
    b = ga.align_sw_matrix_batch(reqs, m)
    assert b.ok(), b.error
    for (t, p), r in zip(pairs, b.results):
        sc, si, sj, ei, ej, cg = oracle.sw_trace(t, p, tab, 11, 1)
        assert (r.score, r.start_i, r.start_j, r.end_i, r.end_j, r.cigar) == (sc, si, sj, ei, ej, cg), (t, p)
        if sc > 0:
            assert r.rescore_ok and r.wellformed_ok
    for r in reqs:
        r.with_cigar = False
    bs = ga.align_sw_matrix_batch(reqs, m)
    assert bs.ok()
    for (t, p), r in zip(pairs, bs.results):
        assert (r.score, r.end_i, r.end_j) == oracle.sw_score_lane(t, p, tab, 11, 1), (t, p)


def test_protein_sw_long_queries(gpu, oracle):
    """Queries of 10..16 strips (several rounds of the multi-wave pipeline) over
    short texts (pipeline period set by the wave lag) and long ones (set by the
    text length): trace and score-only results equal the oracle's."""
    import genoaligner_amd as ga
    m = ga.blosum62(11, 1)
    alpha = m.alphabet.decode()
    S = [[m.score[a][b] for b in range(20)] for a in range(20)]
    tab = oracle.matrix_table(alpha, S, -1)
    rng = random.Random(43)
    pairs = []
    for k in range(24):
        p = _protein(rng, rng.randint(600, 1000))
        if k % 2:
            a = rng.randint(0, len(p) - 200)
            t = _protein(rng, rng.randint(0, 400)) + _mutate(rng, p[a: a + 200], 0.2, alpha) + \
                _protein(rng, rng.randint(0, 800))
        else:
            t = _protein(rng, rng.randint(20, 120))
        pairs.append((t, p))
    for cigar in (True, False):
        reqs = [ga.SWRequest(text=t, text_len=len(t), pattern=p, pattern_len=len(p), with_cigar=cigar)
                for t, p in pairs]
        b = ga.align_sw_matrix_batch(reqs, m)
        assert b.ok(), b.error
        for (t, p), r in zip(pairs, b.results):
            if cigar:
                assert (r.score, r.start_i, r.start_j, r.end_i, r.end_j, r.cigar) == \
                       oracle.sw_trace(t, p, tab, 11, 1), (len(t), len(p))
            else:
                assert (r.score, r.end_i, r.end_j) == oracle.sw_score_lane(t, p, tab, 11, 1), (len(t), len(p))


def _sw_scheme(scores, unknown, go, ge, alphabet="ARNDCQEGHILKMFPSTWYV"):
    import genoaligner_amd._native as N
    m = N.SwMatrix()
    m.alphabet = alphabet.encode()
    for a in range(len(alphabet)):
        for b in range(len(alphabet)):
            m.score[a][b] = scores[a][b]
    m.unknown, m.gap_open, m.gap_extend = unknown, go, ge
    return m


def test_protein_sw_packed_reruns_and_fallbacks(gpu, oracle):
    """Edges of the packed substitution-matrix kernel (two pairs per
    workgroup): a pair whose score leaves its 16-bit range (3,100 residues,
    mostly W, re-run by the int32 kernel) sharing workgroups with ordinary
    pairs; a text past its 4,096 rows (int32 kernel) in the same batch; empty
    sides; a scheme whose lowest score is below -gap_open (the extra bias
    term); and one whose biased scores leave a byte (the whole batch on the
    int32 kernel).  Trace and score-only results equal the oracle's."""
    import genoaligner_amd as ga
    alpha = "ARNDCQEGHILKMFPSTWYV"
    b62 = ga.blosum62(11, 1)
    S62 = [[b62.score[a][b] for b in range(20)] for a in range(20)]
    Sgk = [[(5 if a == b else -((a * 7 + b * 3) % 10)) for b in range(20)] for a in range(20)]
    Swide = [[(120 if a == b else -20) for b in range(20)] for a in range(20)]
    cases = [(b62, S62, -1, 11, 1), (_sw_scheme(Sgk, -3, 4, 1), Sgk, -3, 4, 1),
             (_sw_scheme(Swide, -5, 150, 10), Swide, -5, 150, 10)]
    rng = random.Random(47)
    for m, S, unk, go, ge in cases:
        tab = oracle.matrix_table(alpha, S, unk)
        pairs = []
        for k in range(40):
            p = _protein(rng, rng.randint(1, 500))
            t = _mutate(rng, p, 0.15, alpha) if k % 2 else _protein(rng, rng.randint(1, 700))
            pairs.append((t, p))
        big = "".join(rng.choice("WWWWWWWWWC") for _ in range(3100))
        pairs.insert(5, (big, _mutate(rng, big, 0.002, alpha)))
        longt = _protein(rng, 4500)
        pairs.insert(11, (longt, _mutate(rng, longt[1000:1400], 0.1, alpha)))
        pairs.insert(20, ("", "ACDE"))
        pairs.append(("KLM", ""))
        for cigar in (True, False):
            reqs = [ga.SWRequest(text=t, text_len=len(t), pattern=p, pattern_len=len(p), with_cigar=cigar)
                    for t, p in pairs]
            b = ga.align_sw_matrix_batch(reqs, m)
            assert b.ok(), b.error
            for (t, p), r in zip(pairs, b.results):
                if cigar:
                    assert (r.score, r.start_i, r.start_j, r.end_i, r.end_j, r.cigar) == \
                           oracle.sw_trace(t, p, tab, go, ge), (go, len(t), len(p))
                else:
                    assert (r.score, r.end_i, r.end_j) == oracle.sw_score_lane(t, p, tab, go, ge), \
                           (go, len(t), len(p))
        if m is b62:
            assert oracle.sw_trace(pairs[5][0], pairs[5][1], tab, go, ge)[0] > 32767 - 255 - 11


def test_wfa_graph_replays_match_eager(gpu, oracle):
    """The device-planned submission captured into a HIP graph: every replay
    (outputs reset between replays, eager work of the program in between)
    equals the eager run, which equals the oracle."""
    import torch
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_mixed(n=900, lo=30, hi=3000, seed=404)
    want = (np.random.default_rng(9).random(ps.n) < 0.7).astype(np.uint8)
    eager = WfaDeviceBatch(ps, trace=True, want=want, classes=False)
    eager.run()
    esc, ecg = eager.results()
    for i in range(0, ps.n, 20):
        sc, cg = oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i]))
        assert (int(esc[i]), ecg[i].decode()) == (sc, cg if want[i] else "")
    gb = WfaDeviceBatch(ps, trace=True, want=want, classes=False)
    g = gb.graph()
    other = WfaDeviceBatch(W.wfa_plain(n=500, seed=3), trace=True)
    for rep in range(4):
        gb.score.fill_(-7)
        gb.cig_len.fill_(-7)
        gb.cigar.fill_(0)
        torch.cuda.synchronize()
        g.launch()
        other.run()                                   # the program's own eager work in between
        torch.cuda.synchronize()
        sc, cg = gb.results()
        assert [int(x) for x in sc] == [int(x) for x in esc], rep
        assert cg == ecg, rep
    g.close()


def test_wfa_outputs_invariant_across_tunings(gpu):
    """Every waves-per-workgroup choice of every size class (and the
    autotuner's pick) gives the default configuration's results exactly."""
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch, wfa_autotune, wfa_get_tuning, wfa_set_tuning
    ps = W.wfa_mixed(n=700, lo=40, hi=4000, seed=515)
    want = (np.random.default_rng(3).random(ps.n) < 0.7).astype(np.uint8)
    wfa_set_tuning(None)
    ref = WfaDeviceBatch(ps, trace=True, want=want)
    ref.run()
    rsc, rcg = ref.results()
    try:
        for w in (1, 2, 4, 8):
            wfa_set_tuning([w] * 4)
            assert wfa_get_tuning() == [w] * 4
            for classes in (True, False):
                b = WfaDeviceBatch(ps, trace=True, want=want, classes=classes)
                b.run()
                sc, cg = b.results()
                assert [int(x) for x in sc] == [int(x) for x in rsc], (w, classes)
                assert cg == rcg, (w, classes)
        wfa_set_tuning(None)
        tb = WfaDeviceBatch(ps, trace=True, want=want)
        waves, ms = wfa_autotune(tb, reps=1)
        assert all(w in (1, 2, 4, 8) for w in waves)
        assert wfa_get_tuning() == waves
        b = WfaDeviceBatch(ps, trace=True, want=want)
        b.run()
        sc, cg = b.results()
        assert [int(x) for x in sc] == [int(x) for x in rsc] and cg == rcg
        sc_t, cg_t = tb.results()                 # the tuner's own runs wrote the same results
        assert [int(x) for x in sc_t] == [int(x) for x in rsc] and cg_t == rcg
    finally:
        wfa_set_tuning(None)


def test_wfa_device_plan_large_batch(gpu, oracle):
    """Above the single-workgroup plan size the device plan runs as four
    kernels; its results equal the host-planned run and the oracle."""
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_mixed(n=70000, lo=20, hi=700, seed=616)
    host = WfaDeviceBatch(ps, trace=True, classes=True)
    host.run()
    dev = WfaDeviceBatch(ps, trace=True, classes=False)
    dev.run()
    hs, hc = host.results()
    ds, dc = dev.results()
    assert [int(x) for x in hs] == [int(x) for x in ds]
    assert hc == dc
    for i in range(0, ps.n, 997):
        sc, cg = oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i]))
        assert (int(ds[i]), dc[i].decode()) == (sc, cg)


def test_workspace_blocks_reused_across_streams(gpu):
    """Large workspaces (wavefront stacks) are cached device blocks handed from
    one call to the next; two batches, each on a stream of its own and enqueued
    in turn with no host wait, take each other's blocks and still give the
    results of calls run one at a time."""
    import torch
    from genoaligner_amd import _native as N
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    batches = [WfaDeviceBatch(W.wfa_mixed(n=3000, lo=40, hi=3000, seed=s), trace=True, classes=c)
               for s, c in ((717, True), (718, False))]
    refs = []
    for b in batches:
        b.run()
        torch.cuda.synchronize()
        refs.append(b.results())
    N.lib().ga_trim_workspace()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for _ in range(4):
        for b, st in zip(batches, streams):
            b.run(st)
    torch.cuda.synchronize()
    for b, (rs, rc) in zip(batches, refs):
        sc, cg = b.results()
        assert [int(x) for x in sc] == [int(x) for x in rs]
        assert cg == rc
    N.lib().ga_trim_workspace()


def test_wfa_one_class_one_length_bin(gpu, oracle):
    """A batch of one size class whose pairs all fall in one length bin keeps
    its own order (no claim-order upload): 150 bp pairs at smax 64 through
    the host entry point and the device entry point with host-planned
    classes, against the oracle; plus the same pairs with one long pair
    appended (a second bin, so a sorted claim order)."""
    import genoaligner_amd as ga
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_plain(n=3000, length=150, rate=0.02, smax=64, seed=5)
    exp = [oracle.wfa(ps.a_seq(i), ps.b_seq(i), 64) for i in range(ps.n)]
    b = WfaDeviceBatch(ps, trace=True)
    b.run()
    sc, cg = b.results()
    assert [(int(sc[i]), cg[i].decode()) for i in range(ps.n)] == exp
    reqs = [_areq(ga, ps.a_seq(i).decode(), ps.b_seq(i).decode(), 64) for i in range(300)]
    reqs.append(_areq(ga, "ACGT" * 300, "ACGT" * 290 + "AGGT" * 10, 64))
    r = ga.align_batch_ext(reqs)
    assert r.ok()
    assert [(x.score, x.cigar) for x in r.results[:300]] == exp[:300]
    assert (r.results[300].score, r.results[300].cigar) == oracle.wfa("ACGT" * 300, "ACGT" * 290 + "AGGT" * 10, 64)


def test_sw_validate_device_matches_host_checks(gpu, oracle):
    """ga_sw_validate_device (the CIGAR checks ga_align_sw_batch now runs on
    the device) against the host's ga_sw_validate_cigar, on a device batch's
    own results and on mutations of them written back into its buffers: ops
    flipped, inserted, dropped or foreign, ends and starts moved, scores
    changed, CIGARs running past both sequences."""
    import ctypes as C
    import torch
    from genoaligner_amd import _native as N
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import SwDeviceBatch
    sc = (2, -3, 3, 1)
    ps = W.sw_embedded(n=600)
    dev = SwDeviceBatch(ps, scoring=sc, trace=True)
    dev.run()
    out, cig = dev.results()
    flags = dev.validate()
    assert all(f == (3 if out[k, 0] > 0 else 0) for k, f in enumerate(flags))
    rng = random.Random(77)
    L = N.lib()
    scc = N.SWScoringC(*sc)
    out2 = out.copy()
    raw = bytearray(dev.cigar.cpu().numpy().tobytes())
    lens = np.zeros(ps.n, np.int32)
    for k in range(ps.n):
        c = list(cig[k].decode())
        score, si, sj, ei, ej = (int(x) for x in out[k])
        kind = rng.randrange(9)
        if kind == 1 and c:
            c[rng.randrange(len(c))] = rng.choice("MXID")
        elif kind == 2:
            c.insert(rng.randint(0, len(c)), rng.choice("MXID"))
        elif kind == 3 and c:
            del c[rng.randrange(len(c))]
        elif kind == 4 and c:
            c[rng.randrange(len(c))] = rng.choice("Z=m")
        elif kind == 5:
            ei, ej = ei + rng.randint(-2, 2), ej + rng.randint(-2, 2)
        elif kind == 6:
            si, sj = si + rng.randint(0, 2), sj + rng.randint(0, 2)
        elif kind == 7:
            score += rng.choice([-1, 1])
        elif kind == 8:
            c += ["M"] * rng.randint(1, 20)
        cg = "".join(c)[: int(ps.a_len[k] + ps.b_len[k])]
        end = int(dev.slot_off_np[k] + dev.slot[k])
        raw[end - len(cg): end] = cg.encode("latin-1")
        lens[k] = len(cg)
        out2[k] = (score, si, sj, ei, ej)
    dev.cigar.copy_(torch.frombuffer(raw, dtype=torch.uint8).to(dev.cigar.device))
    dev.out.copy_(torch.from_numpy(out2.reshape(-1).astype(np.int32)).to(dev.out.device))
    dev.cig_len.copy_(torch.from_numpy(lens).to(dev.cig_len.device))
    got = dev.validate()
    r, w = C.c_int32(), C.c_int32()
    seen = set()
    for k in range(ps.n):
        score, si, sj, ei, ej = (int(x) for x in out2[k])
        if score <= 0 or lens[k] == 0:
            assert got[k] == 0, k
            continue
        end = int(dev.slot_off_np[k] + dev.slot[k])
        cg = bytes(raw[end - lens[k]: end])
        t, p = (x if isinstance(x, bytes) else x.encode() for x in (ps.a_seq(k), ps.b_seq(k)))
        assert L.ga_sw_validate_cigar(cg, len(cg), t, len(t), si, ei, p, len(p), sj, ej, C.byref(scc), score,
                                      C.byref(r), C.byref(w)) == 0
        assert got[k] == (r.value | (w.value << 1)), (k, cg, si, ei, sj, ej, score)
        seen.add(int(got[k]))
    assert {0, 2, 3} <= seen and 1 not in seen          # rescore_ok implies wellformed_ok
