"""The MSA surface without a GPU: profiles built by the library's host code
(ga_msa_profile_from_rows) equal the oracle's (pinned to the reference), the
alignment entry point refuses bad batches before any device work, and
merge_rows restates merge_profiles' rows."""
import random

import numpy as np
import pytest

from conftest import read_tsv


def test_profiles_equal_the_oracle(oracle):
    from genoaligner_amd import msa as M
    rng = random.Random(9)
    for alpha, letters in ((4, "ACGTRYMKSWHBVDNXacgtnu-.*"), (20, "ARNDCQEGHILKMFPSTWYVBZJUXOacdx-*.")):
        for _ in range(60):
            L, n = rng.randint(1, 60), rng.randint(1, 8)
            rows = ["".join(rng.choice(letters) for _ in range(L)) for _ in range(n)]
            p = M.profile_from_rows(rows, alpha)
            c, o = oracle.msa_profile(rows, alpha)
            assert np.array_equal(p.cols, c) and np.array_equal(p.occ, o)


def test_align_refusals_before_device_work():
    from genoaligner_amd import msa as M
    a = M.profile_from_seq("ACGT")
    with pytest.raises(M.MsaError, match="alphabet must be 4"):
        M.align_profiles(a, a, M.Params(alpha=7))
    with pytest.raises(M.MsaError, match="equal length"):
        M.profile_from_rows(["ACG", "AC"])
    assert M.align_profiles_batch([], M.Params()) == []


def test_merge_rows_restates_merge_profiles(oracle):
    """Against the oracle's restatement on the golden alignments: the merged
    rows hold every input row's letters in order and are all of the width
    ai + aj + len(cigar) + the two suffixes."""
    from genoaligner_amd import msa as M
    for r in read_tsv("msa_pp.tsv")[:120]:
        A, B = r[1].split(","), r[2].split(",")
        span = tuple(int(x) for x in r[4:8])
        cg = r[8]
        aln = M.AlignResult(0.0, cg, *span)
        pa, pb = M.Profile(None, np.zeros(len(A[0])), A), M.Profile(None, np.zeros(len(B[0])), B)
        got = M.merge_rows(pa, pb, aln)
        assert got == oracle.msa_merge_rows(A, B, span, cg)
        w = span[0] + span[1] + len(cg) + (len(A[0]) - span[2]) + (len(B[0]) - span[3])
        assert all(len(x) == w for x in got)
        assert [x.replace("-", "") for x in got] == [x.replace("-", "") for x in A + B]
