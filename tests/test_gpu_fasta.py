"""GPU tests of the device FASTA ingest (ga_fasta_device): per record the same
sequence bytes as the oracle's reader (a restatement of io::read_fasta,
include/genoaligner/io/fasta.hpp:59-157, pinned in tests/test_fasta.py), the
A/C/G/T flag and 2-bit codes as the oracle computes them, and for malformed
files the reader's message.  Bit-exact."""
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN
from test_fasta import REF_CASES, RULE_CASES, _random_fasta

pytestmark = pytest.mark.gpu


def _check(oracle, data: bytes, shift: int = 0):
    """shift > 0 places the file that many bytes into a device buffer (an
    unaligned start: the byte-load path of the kernels)."""
    import torch
    from genoaligner_amd.fasta import FastaError, fasta_to_device
    recs, err = oracle.fasta_parse(data)
    src = data
    if shift:
        buf = torch.zeros(len(data) + shift, dtype=torch.uint8, device="cuda")
        if data:
            buf[shift:] = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
        src = buf[shift:]
        assert src.data_ptr() % 16 == shift % 16
    try:
        d = fasta_to_device(src)
    except FastaError as e:
        assert err == str(e), (data[:200], err, str(e))
        return None
    assert err is None, (data[:200], err)
    assert d.n == len(recs)
    seqs = [r[2] for r in recs]
    stream = b"".join(seqs)
    assert d.total == len(stream)
    assert d.seq.cpu().numpy().tobytes() == stream
    lens = d.seq_len.cpu().numpy()
    offs = d.seq_off.cpu().numpy()
    assert lens.tolist() == [len(s) for s in seqs]
    assert offs.tolist() == np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64).tolist()
    flags = d.acgt.cpu().numpy().tolist()
    assert flags == [int(len(s) > 0 and set(s) <= set(b"ACGT")) for s in seqs]
    pk = d.pack2.cpu().numpy().view(np.uint32)
    assert np.array_equal(pk, oracle.fasta_pack2(stream))
    return d


@pytest.mark.parametrize("k", range(len(REF_CASES) + len(RULE_CASES)))
def test_device_ingest_reader_cases(gpu, oracle, k):
    _check(oracle, (REF_CASES + RULE_CASES)[k][0])


def test_device_ingest_random_small(gpu, oracle):
    """Short random files: every state of the reader, errors included."""
    rng = random.Random(17)
    bad = 0
    for it in range(400):
        bad += _check(oracle, _random_fasta(rng, rng.randint(0, 300)), shift=it % 3) is None
    assert 40 < bad < 360


# The device ingest cuts a file into tiles of 512 threads x 32 bytes; the
# line state, header count and kept-byte count cross each seam between them.
FASTA_TILE = 512 * 32


def test_device_ingest_tile_seams(gpu, oracle):
    """Files of two to six 16 KiB tiles with a \\r\\n pair, a header and
    sequence bytes planted at a tile seam (one byte before, at, or after it),
    plus random line breaks, headers and \\r\\n pairs elsewhere."""
    rng = random.Random(3)
    for it in range(48):
        parts = [b">r0 x\n"]
        target = rng.randint(2 * FASTA_TILE + 100, 6 * FASTA_TILE)
        while sum(map(len, parts)) < target:
            parts.append(rng.choice([b"ACGT" * rng.randint(1, 40), b"\n", b"\r\n", b"\r", b"acgn",
                                     b"\n>h%d d\n" % it, b"\r\n>q\r\n"]))
        data = b"".join(parts) + b"\nA"
        k = rng.randint(1, len(data) // FASTA_TILE - 1)
        seam = k * FASTA_TILE + (-1, 0, 1)[it % 3]
        # the planted bytes start one before the seam, so \\r, \\n, '>' and
        # the first sequence byte each land on it in turn as `seam` moves
        for cut in (seam - 1, seam + 7):
            data = data[:cut] + b"\r\n>s\nGG" + data[cut:]
        _check(oracle, data, shift=(0, 1, 7, 16)[it % 4])


def test_device_ingest_reference_data_files(gpu, oracle):
    for name, n in (("mtdna_human.fa", 16569), ("mtdna_chimp.fa", 16554)):
        d = _check(oracle, open(os.path.join(GOLDEN, name), "rb").read())
        assert d.n == 1 and d.total == n


def test_device_ingest_genome_scale(gpu, oracle):
    """A 48 MiB multi-record file (N runs, soft-masking, CRLF records)."""
    from genoaligner_amd import workloads as W
    data = W.fasta_genome(48 << 20, seed=4, rec_lo=5_000, rec_hi=1_500_000).tobytes()
    d = _check(oracle, data)
    assert d.n > 20 and 0 < int(d.acgt.sum()) < d.n
    _check(oracle, data[: 9 << 20] + b"\nA", shift=3)


def test_device_ingest_device_tensor_and_errors(gpu, oracle):
    import torch
    from genoaligner_amd.fasta import FastaError, fasta_to_device
    data = b">a\nAC\n>b x\nGT\n"
    t = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    d = fasta_to_device(t, pack2=False)
    assert d.n == 2 and d.pack2 is None and d.seq.cpu().numpy().tobytes() == b"ACGT"
    # more records than the first guess of the wrapper (16): the call reports the count, the wrapper retries
    many = b"".join(b">r%d\nA%s\n" % (k, b"C" * (k % 5)) for k in range(300))
    d = _check(oracle, many)
    assert d.n == 300
    big = b">a\nAC\n" * 3000 + b">b\n"
    with pytest.raises(FastaError) as ei:
        fasta_to_device(big)
    assert str(ei.value) == oracle.fasta_parse(big)[1]
