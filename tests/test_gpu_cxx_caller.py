"""A C++ caller of include/genoaligner/api.hpp (tests/cxx/golden_caller.cpp,
in the shape of the reference's examples/align_fasta.cpp and
examples/align_sw.cpp) linked against the built library: every golden vector
through align_batch / align / align_sw_batch / align_sw, and the two examples'
requests on the human mtDNA record checked against the oracle."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CALLER = os.path.join(ROOT, "tests", "cxx", "_build", "golden_caller")


def _run():
    if not os.path.exists(CALLER):
        pytest.fail("tests/cxx/_build/golden_caller is not built (run __graft_entry__.build())")
    return subprocess.run([CALLER, GOLDEN], capture_output=True, text=True, timeout=300)


def test_caller_links_and_reports_no_device():
    """CPU: the program loads the library and its C++ surface, and says so
    when no device is visible."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    r = _run()
    assert r.returncode == 3 and r.stdout.strip() == "no device"


@pytest.mark.gpu
def test_caller_golden_vectors_and_examples(gpu, oracle):
    r = _run()
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    tally = [ln for ln in lines if ln.startswith("GOLDEN ")]
    assert tally and tally[0].endswith(" bad 0")
    assert int(tally[0].split()[2]) > 2000
    ex = [ln.split(" ") for ln in lines if ln.startswith("EX ")]
    assert len(ex) == 6
    for f in ex:
        if f[1] == "wfa":
            _, _, _, p, t, score, resolved, cigar, rs, wf = f
            assert (int(score), cigar) == oracle.wfa(p, t, 64)
            assert int(score) == 1 and resolved == rs == wf == "1"
        else:
            _, _, _, t, p, score, si, sj, ei, ej, cigar, rs, wf = f
            assert (int(score), int(si), int(sj), int(ei), int(ej), cigar) == \
                oracle.sw_trace(t, p, (2, -3, 3, 1))
            assert rs == wf == "1"


MSA_CALLER = os.path.join(ROOT, "tests", "cxx", "_build", "msa_caller")


def test_msa_caller_links_and_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    if not os.path.exists(MSA_CALLER):
        pytest.fail("tests/cxx/_build/msa_caller is not built (run __graft_entry__.build())")
    r = subprocess.run([MSA_CALLER, GOLDEN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and r.stdout.strip() == "no device"


@pytest.mark.gpu
def test_msa_caller_golden_vectors(gpu, oracle):
    """include/genoaligner/msa/msa.hpp over the MSA golden vectors, then a
    four-sequence progressive merge checked row by row against the oracle."""
    r = subprocess.run([MSA_CALLER, GOLDEN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    assert lines[-1].startswith("GOLDEN checked") and lines[-1].endswith(" bad 0")
    rows = [ln[4:] for ln in lines if ln.startswith("MSA ")]
    assert len(rows) == 4 and len({len(x) for x in rows}) == 1
    seqs = ["ACGTACGTTTGACCA", "ACGTACGTTGACCA", "ACGAACGTTTGACGA", "TTACGTACGTTTGA"]
    assert sorted(x.replace("-", "") for x in rows) == sorted(seqs)
    # replay the merge with the oracle
    ch = oracle.nj_tree(oracle.kmer_distances(seqs, 3), 4)
    node = [[s] for s in seqs]
    P = oracle.MsaParams()
    for l, rr in ch.tolist():
        A, B = node[l], node[rr]
        sc, span, cg = oracle.msa_align(oracle.msa_profile(A), oracle.msa_profile(B), P)
        node.append(oracle.msa_merge_rows(A, B, span, cg))
    assert rows == node[-1]
