"""The profile-profile Gotoh kernel (ga_msa_align_profiles, ga_msa_pp_device)
against the golden vectors written from the reference's own align_profiles
and msa_pp_trace_kernel body, and against the oracle on random and long
profiles.  Scores are floats compared bit for bit; spans and CIGARs exactly."""
import random

import numpy as np
import pytest

from test_msa_oracle import golden_msa

pytestmark = pytest.mark.gpu


def _params(M, tag):
    return {
        "default": M.Params(),
        "legacy": M.Params(psgp=False),
        "global": M.Params(free_end_gaps=False),
        "protein": M.protein_params(),
        "protein_semi": M.Params(alpha=20, sub=M.BLOSUM62.copy(), gap_open=11.0, gap_extend=1.0,
                                 free_end_gaps=True),
    }[tag]


def _oracle_params(O, P):
    return O.MsaParams(P.match, P.ts, P.tv, P.gap_open, P.gap_extend, P.free_end_gaps, P.psgp, P.psgp_scale,
                       P.psgp_min_open, P.psgp_min_ext, P.alpha, None if P.sub is None else P.sub.copy())


def _f32(x):
    return np.float32(x).tobytes()


def test_golden_vectors(gpu):
    from genoaligner_amd import msa as M
    rows = golden_msa()
    by = {}
    for r in rows:
        by.setdefault(r[0], []).append(r)
    n = 0
    for tag, rs in by.items():
        P = _params(M, tag)
        pairs = [(M.profile_from_rows(A, P.alpha), M.profile_from_rows(B, P.alpha)) for _, A, B, *_ in rs]
        got = M.align_profiles_batch(pairs, P)
        for (t, A, B, score, span, cigar), g in zip(rs, got):
            assert _f32(g.score) == _f32(score), (tag, A, B, g.score, score)
            assert ((g.ai, g.aj, g.bi, g.bj), g.cigar) == (span, cigar), (tag, A, B)
            n += 1
    assert n == len(rows)


def _rand_profile(rng, M, alpha, lo, hi, rows_max):
    letters = ("ACGT" * 8 + "RYNacgt-") if alpha == 4 else ("ARNDCQEGHILKMFPSTWYV" * 3 + "BZXJ-")
    L = rng.randint(lo, hi)
    rows = ["".join(rng.choice(letters) for _ in range(L)) for _ in range(rng.randint(1, rows_max))]
    return M.profile_from_rows(rows, alpha)


@pytest.mark.parametrize("tag", ["default", "legacy", "global", "protein", "protein_semi"])
def test_random_profiles_vs_oracle(gpu, oracle, tag):
    """Random profiles of 1..300 columns (several 64-column strips, ragged
    last strips, gap and ambiguity letters, up to 6 rows) in one batch."""
    from genoaligner_amd import msa as M
    rng = random.Random(hash(tag) & 0xffff)
    P = _params(M, tag)
    OP = _oracle_params(oracle, P)
    pairs = [(_rand_profile(rng, M, P.alpha, 1, 300, 6), _rand_profile(rng, M, P.alpha, 1, 300, 6))
             for _ in range(60)]
    got = M.align_profiles_batch(pairs, P)
    for (a, b), g in zip(pairs, got):
        want = oracle.msa_align((a.cols, a.occ), (b.cols, b.occ), OP)
        assert (_f32(g.score), (g.ai, g.aj, g.bi, g.bj), g.cigar) == (_f32(want[0]), want[1], want[2])


def test_long_related_profiles_and_one_against_many(gpu, oracle):
    """Related DNA profiles of 600..1,500 columns (a progressive merge's
    shape), one profile against many (uploaded once), and empty profiles."""
    from genoaligner_amd import msa as M
    rng = random.Random(21)
    P = M.Params()
    OP = _oracle_params(oracle, P)

    def mut(s):
        out = []
        for c in s:
            r = rng.random()
            if r < 0.04:
                continue
            out.append(rng.choice("ACGT") if r < 0.12 else c)
            if rng.random() < 0.03:
                out.append(rng.choice("ACGT"))
        return "".join(out)

    anc = "".join(rng.choice("ACGT") for _ in range(1200))
    seqs = [mut(anc) for _ in range(6)]
    prof = [M.profile_from_seq(s) for s in seqs]
    ab = M.merge_profiles(prof[0], prof[1], M.align_profiles(prof[0], prof[1], P))
    cd = M.merge_profiles(prof[2], prof[3], M.align_profiles(prof[2], prof[3], P))
    bait = M.profile_from_seq(anc[100:700])
    empty = M.Profile(np.zeros((0, 4), np.float32), np.zeros(0, np.float32), [], 4)
    pairs = [(ab, cd), (cd, ab), (ab, prof[4])] + [(bait, p) for p in prof] + [(p, bait) for p in prof]
    pairs += [(empty, bait), (bait, empty), (empty, empty)]
    for Q, OQ in ((P, OP), (M.Params(free_end_gaps=False), _oracle_params(oracle, M.Params(free_end_gaps=False)))):
        got = M.align_profiles_batch(pairs, Q)
        for (a, b), g in zip(pairs, got):
            want = oracle.msa_align((a.cols, a.occ), (b.cols, b.occ), OQ)
            assert (_f32(g.score), (g.ai, g.aj, g.bi, g.bj), g.cigar) == (_f32(want[0]), want[1], want[2])


def test_device_entry_matches_host_entry(gpu):
    from genoaligner_amd import msa as M
    rng = random.Random(5)
    P = M.Params()
    profs = [_rand_profile(rng, M, 4, 20, 400, 4) for _ in range(12)]
    pairs_idx = [(rng.randrange(12), rng.randrange(12)) for _ in range(40)]
    want = M.align_profiles_batch([(profs[a], profs[b]) for a, b in pairs_idx], P)
    got = M.MsaDeviceBatch(profs, pairs_idx, P)
    got.run()
    assert got.results() == want
