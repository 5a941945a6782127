"""The gene QC tool on the GPU (SW + CIGAR through ga_align_sw_batch, the
native walk): its TSV and summary against the oracle's pipeline, which
tests/test_gene_qc.py pins byte for byte to the reference tool itself
(tools/gene_qc.cpp built from its sources) on the CPU.  Byte-exact."""
import io

import pytest

pytestmark = pytest.mark.gpu


def _want(oracle, seqs, refs, code, kw):
    return oracle.gene_qc_tsv(seqs, refs, code=code, **kw)


@pytest.mark.parametrize("n,codons,code,kw,batch", [
    (300, 380, 2, {}, 0),                                   # CYTB-sized, vertebrate mitochondrial code
    (200, 200, 1, {"min_cov": 0.9}, 0),
    (120, 150, 2, {"scoring": (2, -3, 5, 2)}, 37),          # chunked
])
def test_tool_matches_oracle_pipeline(gpu, oracle, tmp_path, n, codons, code, kw, batch):
    from genoaligner_amd import gene_qc as G
    from genoaligner_amd import workloads as W
    seqs, refs = W.gene_family(n, codons=codons, seed=n + code)
    sp, rp = tmp_path / "s.fa", tmp_path / "r.fa"
    sp.write_bytes(seqs)
    rp.write_bytes(refs)
    want_tsv, want_summary = _want(oracle, seqs, refs, code, kw)
    out, err = io.StringIO(), io.StringIO()
    emit = str(tmp_path / "got.tsv")
    rc = G.run(str(sp), str(rp), code=code, emit=emit, out=out, err=err, batch_size=batch,
               **{k: v for k, v in kw.items()})
    assert rc == 0
    assert open(emit, encoding="latin-1").read() == want_tsv
    assert want_summary in err.getvalue()
    assert out.getvalue().startswith("gene_qc: %d records vs ref bait1 (%d bp, frame 1, bait_stops 1, code %d)"
                                     % (n, 3 * codons + 1, code))


def test_tool_cli_to_stdout(gpu, oracle, tmp_path, capsys):
    from genoaligner_amd import gene_qc as G
    from genoaligner_amd import workloads as W
    seqs, refs = W.gene_family(40, codons=100, seed=2)
    sp, rp = tmp_path / "s.fa", tmp_path / "r.fa"
    sp.write_bytes(seqs)
    rp.write_bytes(refs)
    assert G.main(["gene_qc", str(sp), "--ref", str(rp), "--code", "2"]) == 0
    got = capsys.readouterr().out.split("\n", 1)[1]
    assert got == oracle.gene_qc_tsv(seqs, refs, code=2)[0]
