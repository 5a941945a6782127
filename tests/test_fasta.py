"""FASTA reading on the host: the oracle's restatement pinned against the
reference's own test cases (tests/io/test_fasta.cpp:43-166), its data files
(tests/data/mtdna_{human,chimp}.fa, 16,569 and 16,554 bp per
tests/real/test_real_sequences.cpp:21-22) and, where oracle/_ref is built, the
reference's io::read_fasta itself on random files; then the library's reader
(ga_fasta_parse / ga_fasta_read_file, a host call: no GPU needed) against the
oracle."""
import os
import random

import pytest

from conftest import GOLDEN

# (text, expected records or None, expected error or None): the reference's cases
REF_CASES = [
    (b">seq1 first record\nACGTACGTAC\nGTACGTACGT\nACGT\n>seq2 second\nTTTT\n",
     [(b"seq1", b"first record", b"ACGTACGTACGTACGTACGTACGT"), (b"seq2", b"second", b"TTTT")], None),
    (b">a\r\nACGT\r\nACGT\r\n>b\r\nTT\r\n", [(b"a", b"", b"ACGTACGT"), (b"b", b"", b"TT")], None),
    (b"\n>only_id\n\nACGT\n\n\n>with_desc  some words here\nNNNN\n",
     [(b"only_id", b"", b"ACGT"), (b"with_desc", b"some words here", b"NNNN")], None),
    (b">masked\nacgtNNNRYSWacgt\n", [(b"masked", b"", b"acgtNNNRYSWacgt")], None),
    (b"ACGT\n", None, "expected '>' at start of a record, found 'A' (line 1)"),
    (b">empty\n>next\nACGT\n", None, "record has a header but no sequence (line 1)"),
    (b"", None, "no records found (line 1)"),
    (b">   \nACGT\n", None, "empty identifier in header (line 1)"),
    (b">a\nAC\n>b\nTT\n>c\n", None, "record has a header but no sequence (line 5)"),
]

# further cases whose answer follows from the reader's rules by hand
RULE_CASES = [
    (b">x\rAC\rGT", [(b"x", b"", b"ACGT")], None),                       # lone \r ends a line
    (b">x\tdesc\there \t \nA C\n", [(b"x", b"desc\there", b"A C")], None),  # tab split, trailing blanks cut
    (b"  junk line\n\t>x\n>y\nA", [(b"y", b"", b"A")], None),             # blank-led lines skipped before a record
    (b">x\nAC>GT\n>y z\n>GT", None, "record has a header but no sequence (line 3)"),
    (b">x\nA\n\r\n\r\r>y\nC", [(b"x", b"", b"A"), (b"y", b"", b"C")], None),
    (b"\n\n#c\n>x\nA", None, "expected '>' at start of a record, found '#' (line 3)"),
    (b">x\n \n", [(b"x", b"", b" ")], None),                            # a space line is sequence
    (b">x", None, "record has a header but no sequence (line 1)"),
    (b"\r\n\r\n", None, "no records found (line 3)"),
    (b">x\nA\n>\tq\nC", None, "empty identifier in header (line 3)"),
]


def _random_fasta(rng: random.Random, n: int) -> bytes:
    alpha = [b">", b"\n", b"\r", b"\r\n", b" ", b"\t", b"A", b"C", b"G", b"T", b"N", b"a", b"x", b"\xe9"]
    w = [6, 10, 3, 3, 2, 1, 8, 8, 8, 8, 1, 1, 1, 1]
    out = []
    if rng.random() < 0.8:
        out.append(b">r0 d\n")
    out += rng.choices(alpha, weights=w, k=n)
    return b"".join(out)


@pytest.mark.parametrize("text,recs,err", REF_CASES + RULE_CASES)
def test_oracle_reader_cases(oracle, text, recs, err):
    assert oracle.fasta_parse(text) == (recs, err)


def test_oracle_reads_reference_data_files(oracle):
    for name, rid, n in (("mtdna_human.fa", b"NC_012920.1", 16569), ("mtdna_chimp.fa", b"NC_001643.1", 16554)):
        recs, err = oracle.fasta_parse(open(os.path.join(GOLDEN, name), "rb").read())
        assert err is None and len(recs) == 1
        assert recs[0][0] == rid and len(recs[0][2]) == n
        assert set(recs[0][2]) <= set(b"ACGTN")


def test_oracle_matches_reference_reader(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    for text, recs, err in REF_CASES + RULE_CASES:
        assert oracle.ref_fasta(text) == (recs, err), text
    rng = random.Random(5)
    errors = 0
    for it in range(3000):
        t = _random_fasta(rng, rng.randint(0, 120))
        want = oracle.ref_fasta(t)
        assert oracle.fasta_parse(t) == want, t
        errors += want[0] is None
    assert 300 < errors < 2700          # both outcomes well covered


def _host_read(data):
    from genoaligner_amd.fasta import FastaError, read_fasta
    try:
        return [(r.id, r.description, r.sequence) for r in read_fasta(data)], None
    except FastaError as e:
        return None, str(e)


@pytest.mark.parametrize("text,recs,err", REF_CASES + RULE_CASES)
def test_library_reader_cases(text, recs, err):
    assert _host_read(text) == (recs, err)


def test_library_reader_matches_oracle_random(oracle):
    rng = random.Random(9)
    for it in range(3000):
        t = _random_fasta(rng, rng.randint(0, 200))
        assert _host_read(t) == oracle.fasta_parse(t), t


def test_library_read_file(tmp_path, oracle):
    from genoaligner_amd.fasta import FastaError, read_fasta_file
    p = tmp_path / "x.fa"
    p.write_bytes(b">x\nACGT\n")
    r = read_fasta_file(str(p))
    assert len(r) == 1 and r[0].id == b"x" and r[0].sequence == b"ACGT"
    missing = str(tmp_path / "definitely_not_here_9f3a.fa")
    with pytest.raises(FastaError) as ei:
        read_fasta_file(missing)
    assert str(ei.value) == "cannot open '%s'" % missing
    if oracle.ref_available():
        assert oracle.ref_fasta_file(missing) == (None, str(ei.value))
        assert oracle.ref_fasta_file(str(p)) == (1, None)
    for name in ("mtdna_human.fa", "mtdna_chimp.fa"):
        path = os.path.join(GOLDEN, name)
        got = [(x.id, x.description, x.sequence) for x in read_fasta_file(path)]
        assert (got, None) == oracle.fasta_parse(open(path, "rb").read())


def test_oracle_pack2(oracle):
    import numpy as np
    w = oracle.fasta_pack2(b"ACGT" * 4 + b"TN")
    assert w.dtype == np.uint32 and len(w) == 2
    assert int(w[0]) == sum(((j % 4) << (2 * j)) for j in range(16))
    assert int(w[1]) == 3
    assert len(oracle.fasta_pack2(b"")) == 0


def _big_fasta(rng: random.Random, nbytes: int) -> bytes:
    """A valid file of about nbytes: blank and blank-led lines before the first
    header, headers with and without descriptions, sequence lines ending in
    \\n, \\r\\n, a lone \\r or a blank line, holding spaces, lowercase, N and '>'
    (never first on a line)."""
    out, size, k = [b"\n \t x>y\n\r\n"], 0, 0
    while size < nbytes:
        head = b">r%d desc %d\t \n" % (k, k) if rng.random() < 0.9 else b">r%d\r\n" % k
        lines = []
        for _ in range(rng.randint(1, 40)):
            ln = bytes(rng.choices(b"ACGTNacgt >x", k=rng.randint(0, 120)))
            if ln.startswith(b">"):
                ln = b"A" + ln
            lines.append(ln + rng.choice([b"\n", b"\r\n", b"\r", b"\n\n"]))
        lines.append(b"ACGT\n")
        rec = head + b"".join(lines)
        out.append(rec)
        size += len(rec)
        k += 1
    return b"".join(out)


def test_library_reader_large_files_match_oracle(oracle):
    """Files of 4 MiB and more are parsed in segments cut at record starts on
    the host's worker threads: the same records as the oracle's reader, and
    for a malformed file its message and line number, wherever the fault
    lies."""
    rng = random.Random(21)
    files = [_big_fasta(rng, 6 << 20), _big_fasta(rng, (4 << 20) + 7)]
    # one record larger than every segment; a file led by 5 MiB of blank lines
    files.append(b">only one\n" + b"ACGTN" * (1 << 20) + b"\n")
    files.append(b"\n" * (5 << 20) + b">late\nACGT\r\n>late2 x\rAC")
    base = _big_fasta(rng, 5 << 20)
    for frac in (0.03, 0.5, 0.97):
        cut = base.index(b"\n>", int(len(base) * frac)) + 1
        files.append(base[:cut] + b">empty_record\n" + base[cut:])   # header but no sequence
        files.append(base[:cut] + b">\nACGT\n" + base[cut:])           # empty identifier
        files.append(base[:cut] + b"> x\nACGT\n" + base[cut:])         # id starts with a blank
    files.append(b"\n" * (5 << 20))                                     # no records
    files.append(base + b">tail_without_sequence\n\n")
    for t in files:
        want = oracle.fasta_parse(t)
        assert _host_read(t) == want, (len(t), want[1])
