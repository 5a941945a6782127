"""The gene QC task (reference tools/gene_qc.cpp) without a GPU: the oracle's
walk on hand-worked CIGARs; the oracle's whole pipeline (FASTA reader, bait
choice, SW batch, walk, flags) against the reference tool itself (built from
its sources into oracle/_ref by oracle/build_ref.sh) on generated gene
families; the library's native walk (ga_gene_qc_walk, a host call) against the
oracle's; and the tool's command-line and FASTA errors."""
import ctypes as C
import io

import numpy as np
import pytest


@pytest.mark.parametrize("cig,text,start_j,frame,code,want", [
    # ATG TAA GGG read in frame 0: one stop, not the last codon
    (b"M" * 9, b"ATGTAAGGG", 0, 0, 1, (1.0, 0, 0, 1, 0)),
    # the span's last complete codon is the stop: the gene's own terminator
    (b"M" * 9, b"ATGGGGTAA", 0, 0, 1, (1.0, 0, 0, 0, 1)),
    # AGA is a stop under code 2 only
    (b"M" * 9, b"ATGAGAGGG", 0, 0, 1, (1.0, 0, 0, 0, 0)),
    (b"M" * 9, b"ATGAGAGGG", 0, 0, 2, (1.0, 0, 0, 1, 0)),
    # one inserted base: a non-multiple-of-3 run, net +1; the codon holding it is no codon
    (b"MMMMIMMMMM", b"ATGTxAAGGG", 0, 0, 1, (1.0, 1, 1, 0, 0)),
    # two deleted bait bases: run 2, net -2 = +1 (mod 3, signed); after them the
    # bait codon at columns 6-8 pairs text AAG
    (b"MMMDDMMMM", b"ATGTAAGGG", 0, 0, 1, (1.0, 1, 1, 0, 0)),
    # an in-frame 3-base deletion: no run counted, net 0
    (b"MMMDDDMMM", b"ATGGGG", 0, 0, 1, (1.0, 0, 0, 0, 0)),
    # frame 1 with the span starting at bait column 1: codons TAA, GGG
    (b"M" * 6, b"TAAGGG", 1, 1, 1, (1.0, 0, 0, 1, 0)),
])
def test_oracle_walk_hand_cases(oracle, cig, text, start_j, frame, code, want):
    ops_i = sum(1 for c in cig if c in b"MXI")
    got = oracle.gene_qc_walk(cig, text, 5, True, 0, start_j, ops_i - 1, frame, code)
    assert got[1:] == want[1:] and abs(got[0] - ops_i / len(text)) < 1e-12


def test_oracle_walk_unaligned(oracle):
    assert oracle.gene_qc_walk(b"MMM", b"ATG", 0, True, 0, 0, 2, 0, 1) == (0.0, 0, 0, 0, 0)
    assert oracle.gene_qc_walk(b"MMM", b"ATG", 5, False, 0, 0, 2, 0, 1) == (0.0, 0, 0, 0, 0)


def test_oracle_bait_choice(oracle):
    # frame 0 with ATG and frame 1 with ATG both stop-free: the longer record wins
    bait = b"ATG" + b"GGC" * 20
    assert oracle.gene_qc_bait([b"TAGTAGTAG", bait, b"C" + bait], 1) == (2, 1, 0)
    assert oracle.gene_qc_bait([b"C" + bait, bait], 1) == (0, 1, 0)


@pytest.mark.parametrize("code,args", [(2, ()), (1, ("--min-cov", "0.9", "--min-score", "150")),
                                       (2, ("--scoring", "2,-3,5,2"))])
def test_oracle_pipeline_matches_reference_tool(oracle, tmp_path, code, args):
    if not oracle.ref_available() or not __import__("os").path.exists(oracle.REF_GENE_QC):
        pytest.skip("oracle/_ref not built")
    from genoaligner_amd import workloads as W
    seqs, refs = W.gene_family(60, codons=120, seed=code + len(args))
    sp, rp = tmp_path / "s.fa", tmp_path / "r.fa"
    sp.write_bytes(seqs)
    rp.write_bytes(refs)
    rc, out, err, tsv = oracle.ref_gene_qc(str(sp), str(rp), ("--code", str(code)) + args,
                                           emit=str(tmp_path / "o.tsv"))
    kw = {}
    for k in range(0, len(args), 2):
        if args[k] == "--min-cov":
            kw["min_cov"] = float(args[k + 1])
        elif args[k] == "--min-score":
            kw["min_score"] = int(args[k + 1])
        else:
            kw["scoring"] = tuple(int(x) for x in args[k + 1].split(","))
    want_tsv, want_summary = oracle.gene_qc_tsv(seqs, refs, code=code, **kw)
    assert rc == 0
    assert tsv == want_tsv
    assert want_summary in err
    flags = [ln.split("\t")[-1] for ln in tsv.splitlines()[1:]]
    assert any(f == "OK" for f in flags) and any("FS=" in f for f in flags) and any("STOP=" in f for f in flags)


def test_library_walk_matches_oracle(oracle):
    from genoaligner_amd import _native as N
    from genoaligner_amd import workloads as W
    from genoaligner_amd.api import _SW_RES
    seqs, refs = W.gene_family(50, codons=90, seed=7)
    recs = [r[2] for r in oracle.fasta_parse(seqs)[0]]
    bait = oracle.fasta_parse(refs)[0][1][2]
    rng = np.random.default_rng(3)
    for frame, code in ((0, 1), (1, 2), (2, 2), (0, 5)):
        res = np.zeros(len(recs), dtype=_SW_RES)
        arena = bytearray()
        want = []
        for k, t in enumerate(recs):
            sc, si, sj, ei, ej, cig = oracle.sw_trace(t, bait, (1, -1, 2, 1))
            resolved = bool(rng.random() < 0.9)
            res[k] = (sc, int(resolved), len(arena), len(cig), si, sj, ei, ej, 1, 1, 0)
            arena += cig.encode()
            want.append(oracle.gene_qc_walk(cig.encode(), t, sc, resolved, si, sj, ei, frame, code))
        buf = np.frombuffer(b"".join(recs) + b"\0", np.uint8)
        lens = np.array([len(t) for t in recs], np.int32)
        off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        rows = (N.QcRow * len(recs))()
        ab = np.frombuffer(bytes(arena) + b"\0", np.uint8)
        assert N.lib().ga_gene_qc_walk(res.ctypes.data, len(recs), ab.ctypes.data, buf.ctypes.data,
                                       off.ctypes.data, lens.ctypes.data, frame, code, rows) == 0
        got = [(r.cov_text, r.fs_runs, r.fs_net, r.stops, r.stop_end) for r in rows]
        assert got == want


def test_tool_argument_and_fasta_errors(tmp_path, capsys):
    from genoaligner_amd import gene_qc as G
    assert G.main(["gene_qc", "a.fa"]) == 2
    assert G.main(["gene_qc", "a.fa", "--code", "2"]) == 2
    assert "--ref required" in capsys.readouterr().err
    bad = tmp_path / "bad.fa"
    bad.write_bytes(b">x\n>y\nACGT\n")
    err = io.StringIO()
    assert G.run(str(bad), str(bad), err=err) == 1
    assert err.getvalue() == "fasta error: record has a header but no sequence (line 1)\n"


def test_library_tsv_matches_oracle(oracle):
    """ga_gene_qc_tsv (a host call) over the oracle's SW results and walk rows
    writes the oracle pipeline's TSV byte for byte, with its summary counts,
    under thresholds that produce every flag and flag combination."""
    from genoaligner_amd import _native as N
    from genoaligner_amd import workloads as W
    from genoaligner_amd.api import _SW_RES
    seqs, refs = W.gene_family(80, codons=90, seed=11)
    recs = oracle.fasta_parse(seqs)[0]
    best, frame, _ = oracle.gene_qc_bait([r[2] for r in oracle.fasta_parse(refs)[0]], 2)
    flags = set()
    for min_cov, min_score in ((0.5, 0), (0.95, 100), (0.0, -1), (1.1, 0), (1.1, 10 ** 6)):
        want_tsv, want_summary = oracle.gene_qc_tsv(seqs, refs, code=2, min_cov=min_cov, min_score=min_score)
        bait = oracle.fasta_parse(refs)[0][best][2]
        res = np.zeros(len(recs), dtype=_SW_RES)
        rows = np.zeros(len(recs), dtype=np.dtype([("cov_text", "<f8"), ("fs_runs", "<i4"), ("fs_net", "<i4"),
                                                   ("stops", "<i4"), ("stop_end", "<i4")]))
        for k, r in enumerate(recs):
            sc, si, sj, ei, ej, cig = oracle.sw_trace(r[2], bait, (1, -1, 2, 1))
            res[k] = (sc, 1, 0, len(cig), si, sj, ei, ej, 1, 1, 0)
            rows[k] = oracle.gene_qc_walk(cig.encode(), r[2], sc, True, si, sj, ei, frame, 2)
        ids = np.frombuffer(b"".join(r[0] for r in recs) + b"\0", np.uint8)
        id_off = np.concatenate([[0], np.cumsum([len(r[0]) for r in recs])]).astype(np.int64)
        lens = np.array([len(r[2]) for r in recs], np.int32)
        counts = np.zeros(6, np.int64)
        L = N.lib()
        args = (res.ctypes.data, rows.ctypes.data, len(recs), ids.ctypes.data, id_off.ctypes.data,
                lens.ctypes.data, min_cov, min_score)
        need = L.ga_gene_qc_tsv(*args, None, 0, counts.ctypes.data)
        out = np.zeros(need, np.uint8)
        assert L.ga_gene_qc_tsv(*args, out.ctypes.data, need - 1, counts.ctypes.data) == need
        assert not out.any()                     # too small: nothing written
        assert L.ga_gene_qc_tsv(*args, out.ctypes.data, need, counts.ctypes.data) == need
        assert out.tobytes().decode("latin-1") == want_tsv
        assert want_summary.endswith("OK=%d NO_ALIGN=%d LOW_COV=%d FS=%d STOP=%d | invalid=%d" % tuple(counts))
        flags |= {ln.split("\t")[-1].split("=")[0] for ln in want_tsv.splitlines()[1:]}
    assert {"OK", "NO_ALIGN", "LOW_COV", "FS", "STOP", "LOW_COV+FS", "LOW_COV+STOP"} <= flags, flags
    assert L.ga_gene_qc_tsv(None, None, 3, None, None, None, 0.5, 0, None, 0, counts.ctypes.data) == -1
    header = b"acc\tlen\tscore\tsi\tsj\tei\tej\tcov_text\tfs_runs\tfs_net\tstops\tstop_end\tflag\n"
    assert L.ga_gene_qc_tsv(None, None, 0, None, None, None, 0.5, 0, None, 0, counts.ctypes.data) == len(header)
