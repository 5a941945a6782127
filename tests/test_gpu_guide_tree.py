"""k-mer distances and neighbour joining on the GPU (ga_kmer_distances,
ga_nj_tree) against the golden vectors written from the reference's own
kmer_distances and nj_tree, and against the oracle on larger random sets
(floats and trees compared exactly)."""
import random

import numpy as np
import pytest

from test_msa_oracle import golden_kmer_nj

pytestmark = pytest.mark.gpu


def test_golden_vectors(gpu):
    from genoaligner_amd import msa as M
    cases = golden_kmer_nj()
    for k, alpha, seqs, D, ch in cases:
        got = M.kmer_distances(seqs, k, alpha)
        assert np.array_equal(got, D), (k, alpha, len(seqs))
        if len(seqs) >= 2:
            assert np.array_equal(M.nj_children(D, len(seqs)), ch), (k, alpha, len(seqs))


def _family(rng, n, L, rate, alpha="ACGT"):
    anc = "".join(rng.choice(alpha) for _ in range(L))
    out = []
    for _ in range(n):
        s = []
        for c in anc:
            r = rng.random()
            if r < rate / 3:
                continue
            s.append(rng.choice(alpha) if r < rate else c)
        out.append("".join(s))
    return out


@pytest.mark.parametrize("k,alpha,n", [(5, 4, 300), (8, 4, 120), (3, 4, 64), (2, 20, 200), (3, 20, 80)])
def test_random_sets_vs_oracle(gpu, oracle, k, alpha, n):
    """Families of related sequences (structure for NJ), duplicates (ties),
    ambiguity letters, sequences shorter than k."""
    from genoaligner_amd import msa as M
    rng = random.Random(k * 1000 + n)
    letters = "ACGT" if alpha == 4 else "ARNDCQEGHILKMFPSTWYV"
    seqs = []
    while len(seqs) < n - 4:
        seqs += _family(rng, rng.randint(2, 12), rng.randint(30, 400), rng.uniform(0.02, 0.3), letters)
    seqs = seqs[: n - 4] + [seqs[0], seqs[1], "N" * 20 + letters * 3, letters[:1]]
    D = M.kmer_distances(seqs, k, alpha)
    want = oracle.kmer_distances(seqs, k, alpha)
    assert np.array_equal(D, want)
    assert np.array_equal(M.nj_children(D, n), oracle.nj_tree(want, n))


def test_nj_on_tied_and_integer_distances(gpu, oracle):
    """Many equal q values: integer distances on 150 leaves, and all-equal
    distances (every candidate ties; the first in scan order must win)."""
    from genoaligner_amd import msa as M
    rng = np.random.default_rng(4)
    n = 150
    D = rng.integers(1, 6, size=n * (n - 1) // 2).astype(np.float32)
    assert np.array_equal(M.nj_children(D, n), oracle.nj_tree(D, n))
    E = np.ones(40 * 39 // 2, dtype=np.float32)
    assert np.array_equal(M.nj_children(E, 40), oracle.nj_tree(E, 40))
    t = M.nj_tree(np.array([0.5], dtype=np.float32), 2)
    assert t.nodes[2] == (0, 1) and t.root == 2
