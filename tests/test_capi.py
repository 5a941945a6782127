"""The C-ABI library without a GPU: it loads, exports every function that
include/genoaligner/genoaligner_c.h declares, its struct layouts match the
Python binding, and requests that the reference refuses are refused before any
device work (src/api/api.cpp:245-266, :600-616, :515-536)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "genoaligner", "genoaligner_c.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from genoaligner_amd import _native as N
    L = N.lib()
    declared = _declared_functions()
    assert set(declared) == set(N.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf" T {name}$", nm, re.M), name


def test_struct_layouts_match_header(tmp_path):
    from genoaligner_amd import _native as N
    src = tmp_path / "layout.c"
    checks = {
        "ga_align_req": N.AlignReq, "ga_align_res": N.AlignRes, "ga_sw_scoring": N.SWScoringC,
        "ga_sw_req": N.SWReq, "ga_sw_res": N.SWRes, "ga_wfa_dev_batch": N.WfaDevBatch,
        "ga_sw_dev_batch": N.SwDevBatch, "ga_route_stats": N.RouteStats, "ga_wfa_tuning": N.WfaTuning,
        "ga_fasta_dev_out": N.FastaDevOut, "ga_qc_row": N.QcRow, "ga_msa_params": N.MsaParams,
        "ga_msa_profile": N.MsaProfile, "ga_msa_pp_res": N.MsaPPRes, "ga_msa_dev_batch": N.MsaDevBatch,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in checks.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", str(src), "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in checks.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_backend_name():
    import genoaligner_amd as ga
    assert ga.backend_name() == "rocm"


@pytest.mark.parametrize("mutate,msg", [
    (lambda r: setattr(r, "smax", 512), "smax out of range [0, 511]"),
    (lambda r: setattr(r, "smax", -1), "smax out of range [0, 511]"),
    (lambda r: setattr(r, "pattern_len", -1), "negative length"),
    (lambda r: setattr(r, "pattern", None), "null pointer with non-zero length"),
])
def test_wfa_refusals_before_device_work(mutate, msg):
    import genoaligner_amd as ga
    good = ga.AlignRequest(pattern="ACGT", pattern_len=4, text="ACGA", text_len=4, smax=8)
    bad = ga.AlignRequest(pattern="ACGT", pattern_len=4, text="ACGA", text_len=4, smax=8)
    mutate(bad)
    b = ga.align_batch([good, bad])
    assert b.status == ga.Status.invalid_argument and b.error == msg
    assert all(r.score == -1 and not r.resolved for r in b.results)
    assert b.resolved_count == 0 and b.unresolved_count == 0


def test_wfa_ext_bound():
    import genoaligner_amd as ga
    b = ga.align_batch_ext([ga.AlignRequest(pattern="A", pattern_len=1, text="A", text_len=1, smax=4001)])
    assert b.status == ga.Status.invalid_argument


@pytest.mark.parametrize("scoring,with_cigar,msg", [
    ((0, -3, 5, 2), True, "SW scoring: match must be > 0"),
    ((2, 2, 5, 2), True, "SW scoring: mismatch must be < match"),
    ((2, -3, 0, 2), True, "SW scoring: gap costs must be > 0"),
    ((2, -3, 5, 0), True, "SW scoring: gap costs must be > 0"),
    ((2, -3, 1, 2), True, "SW scoring: gap_extend > gap_open with with_cigar"),
])
def test_sw_refusals_before_device_work(scoring, with_cigar, msg):
    import genoaligner_amd as ga
    r = ga.SWRequest(text="ACGT", text_len=4, pattern="ACGT", pattern_len=4,
                     scoring=ga.SWScoring(*scoring), with_cigar=with_cigar)
    b = ga.align_sw_batch([r])
    assert b.status == ga.Status.invalid_argument and b.error.startswith(msg)


def test_sw_mixed_schemes_refused():
    import genoaligner_amd as ga
    a = ga.SWRequest(text="ACGT", text_len=4, pattern="ACGT", pattern_len=4)
    b = ga.SWRequest(text="ACGT", text_len=4, pattern="ACGT", pattern_len=4,
                     scoring=ga.SWScoring(2, -3, 5, 1))
    out = ga.align_sw_batch([a, b])
    assert out.status == ga.Status.invalid_argument
    assert out.error == "all requests in an SW batch must share one scoring scheme"


def test_empty_batches():
    import genoaligner_amd as ga
    assert ga.align_batch([]).ok() and ga.align_sw_batch([]).ok()


needs_ref = pytest.mark.skipif("not __import__('oracle.oracle', fromlist=['x']).ref_available()",
                               reason="reference shim library not built")


@needs_ref
def test_refusals_match_reference():
    """The same malformed batches the reference API refuses are refused here."""
    import numpy as np
    from oracle import oracle as O
    import genoaligner_amd as ga
    R = O.ref()
    cases = [([4, 4], [4, 4], [0, 0], [0, 0], [8, 8], 0), ([4, 4], [4, 4], [0, 0], [0, 0], [8, 600], 1),
             ([4, -1], [4, 4], [0, 0], [0, 0], [8, 8], 1), ([4, 4], [4, 4], [0, 1], [0, 0], [8, 8], 1)]
    for pl, tl, npat, ntxt, smax, want in cases:
        a = [np.array(x, dtype=np.int32) for x in (pl, tl, smax)]
        nb = [np.array(x, dtype=np.uint8) for x in (npat, ntxt)]
        st = R.ga_ref_align_batch_status(a[0].ctypes.data, a[1].ctypes.data, nb[0].ctypes.data,
                                         nb[1].ctypes.data, a[2].ctypes.data, 2)
        reqs = [ga.AlignRequest(pattern=None if npat[k] else "A" * max(pl[k], 0), pattern_len=pl[k],
                                text=None if ntxt[k] else "A" * max(tl[k], 0), text_len=tl[k],
                                smax=smax[k]) for k in range(2)]
        if want == 1:
            assert st == 1 and ga.align_batch(reqs).status == ga.Status.invalid_argument
        else:
            assert st == 0


def test_router_set_refusals_and_unload():
    """ga_router_set: a quantile above 1 is refused before any device work;
    NULL unloads (no device needed)."""
    import genoaligner_amd as ga
    from genoaligner_amd import _native as N
    L = N.lib()
    err = C.c_char_p()
    params = np.zeros(N.GA_MLP_PARAMS, np.float32)
    assert L.ga_router_set(params.ctypes.data, 1.5, C.byref(err)) == N.GA_INVALID_ARGUMENT
    assert b"quantile" in err.value
    assert L.ga_router_set(None, 0.95, C.byref(err)) == N.GA_OK
    assert L.ga_router_enabled() == 0
    with pytest.raises(ValueError):
        ga.set_router(np.zeros(10, np.float32))


def test_blosum62_table():
    """The library's BLOSUM62 equals the reference's table
    (src/msa/msa_ref.cpp:116-137): the published matrix except row M, column H,
    which the reference holds as -1 (published: -2), so the table is symmetric
    everywhere else."""
    import genoaligner_amd as ga
    m = ga.blosum62(11, 1)
    alpha = m.alphabet.decode()
    assert alpha == "ARNDCQEGHILKMFPSTWYV"
    S = np.array([[m.score[a][b] for b in range(20)] for a in range(20)])
    ix = {c: i for i, c in enumerate(alpha)}
    asym = {(alpha[i], alpha[j]) for i in range(20) for j in range(20) if S[i, j] != S[j, i]}
    assert asym == {("M", "H"), ("H", "M")}
    assert (S[ix["M"], ix["H"]], S[ix["H"], ix["M"]]) == (-1, -2)
    diag = dict(zip(alpha, np.diag(S)))
    assert diag == {"A": 4, "R": 5, "N": 6, "D": 6, "C": 9, "Q": 5, "E": 5, "G": 6, "H": 8, "I": 4,
                    "L": 4, "K": 5, "M": 5, "F": 6, "P": 7, "S": 4, "T": 5, "W": 11, "Y": 7, "V": 4}
    for a, b, v in (("W", "C", -2), ("Y", "F", 3), ("I", "V", 3), ("D", "E", 2), ("W", "P", -4), ("K", "R", 2)):
        assert S[ix[a], ix[b]] == v
    assert (m.unknown, m.gap_open, m.gap_extend) == (-1, 11, 1)


@pytest.mark.parametrize("mutate,msg", [
    (lambda m: setattr(m, "alphabet", b""), "alphabet"),
    (lambda m: setattr(m, "alphabet", b"ARNA"), "repeated"),
    (lambda m: setattr(m, "gap_open", -1), "gap costs"),
])
def test_matrix_refusals(mutate, msg):
    import genoaligner_amd as ga
    m = ga.blosum62(11, 1)
    mutate(m)
    r = ga.SWRequest(text="ARND", text_len=4, pattern="ARN", pattern_len=3, with_cigar=True)
    b = ga.align_sw_matrix_batch([r], m)
    assert b.status == ga.Status.invalid_argument and msg in b.error


def test_matrix_gap_rule_refused_with_cigar_only():
    import genoaligner_amd as ga
    m = ga.blosum62(1, 3)
    r = ga.SWRequest(text="ARND", text_len=4, pattern="ARN", pattern_len=3, with_cigar=True)
    b = ga.align_sw_matrix_batch([r], m)
    assert b.status == ga.Status.invalid_argument and "gap_extend > gap_open" in b.error


def test_cpp_surface_compiles_links_and_refuses(tmp_path):
    """A C++ caller of include/genoaligner/api.hpp builds against the library
    and sees the reference's refusals (no device work is needed for them)."""
    from genoaligner_amd import _native as N
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "genoaligner/api.hpp"
#include <cstdio>
int main() {
    genoaligner::AlignRequest r;
    r.pattern = "ACGT"; r.pattern_len = 4; r.text = "ACGA"; r.text_len = 4; r.smax = 600;
    auto b = genoaligner::align_batch({r});
    std::printf("wfa %d %s\n", int(b.status), b.error ? b.error : "-");
    auto m = genoaligner::blosum62(11, 1);
    std::printf("b62 %s %d %d %d\n", m.alphabet.c_str(), m.score[17][17], m.score[12][8], m.gap_open);
    m.alphabet = "";
    genoaligner::SWRequest q;
    q.text = "ARND"; q.text_len = 4; q.pattern = "ARN"; q.pattern_len = 3;
    auto s = genoaligner::align_sw_matrix_batch({q}, m);
    std::printf("sw %d %s\n", int(s.status), s.error ? s.error : "-");
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(N.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", str(src), "-I", os.path.join(ROOT, "include"), "-L", libdir,
                           "-lgenoaligner_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).splitlines()
    assert out[0] == "wfa 1 smax out of range [0, 511]"
    assert out[1] == "b62 ARNDCQEGHILKMFPSTWYV 11 -1 11"
    assert out[2].startswith("sw 1 SW matrix: alphabet")


def test_request_marshalling():
    """The Python layer's request arrays: the C structs' layout, each sequence
    at its address followed by zero bytes up to its declared length, null for
    None, str / bytes / bytearray alike."""
    import ctypes as C
    import numpy as np
    from genoaligner_amd import api as A, _native as N
    for cs, dt in ((N.AlignReq, A._ALIGN_REQ), (N.AlignRes, A._ALIGN_RES),
                   (N.SWReq, A._SW_REQ), (N.SWRes, A._SW_RES)):
        assert dt.itemsize == C.sizeof(cs)
        for name, _ in cs._fields_:
            assert dt.fields[name][1] == getattr(cs, name).offset, (cs.__name__, name)
    keep = []
    seqs = ["ACGT", None, b"GG", bytearray(b"T"), ""]
    lens = [4, 0, 5, 1, 0]
    addr = A._seq_table(seqs, lens, keep)
    assert addr[1] == 0 and all(addr[i] != 0 for i in (0, 2, 3, 4))
    assert C.string_at(int(addr[0]), 5) == b"ACGT\0"
    assert C.string_at(int(addr[2]), 6) == b"GG\0\0\0\0"
    assert C.string_at(int(addr[3]), 2) == b"T\0"
    assert C.string_at(int(addr[4]), 1) == b"\0"
    arr = np.zeros(2, dtype=A._SW_REQ)
    arr["scoring"] = [(2, -3, 5, 2), (1, -1, 2, 1)]
    arr["with_cigar"] = [1, 0]
    p = C.cast(A._ptr(arr, N.SWReq), C.POINTER(N.SWReq))
    assert (p[1].scoring.match, p[1].scoring.mismatch, p[1].scoring.gap_open, p[1].scoring.gap_extend,
            p[1].with_cigar) == (1, -1, 2, 1, 0)


def test_cpp_fasta_surface(tmp_path):
    """A C++ caller of include/genoaligner/io/fasta.hpp (the reference's
    io::read_fasta / read_fasta_file surface) gets the reference's records and
    messages (tests/io/test_fasta.cpp:43-166 cases); host only."""
    from genoaligner_amd import _native as N
    src = tmp_path / "fa.cpp"
    src.write_text(r'''
#include "genoaligner/io/fasta.hpp"
#include <cstdio>
using namespace genoaligner::io;
static void show(const char* text) {
    FILE* f = std::tmpfile();
    std::fputs(text, f);
    std::rewind(f);
    std::vector<FastaRecord> r;
    std::string err;
    const bool ok = read_fasta(f, &r, &err);
    std::fclose(f);
    std::printf("%d %zu", int(ok), r.size());
    for (auto& x : r) std::printf(" [%s|%s|%s]", x.id.c_str(), x.description.c_str(), x.sequence.c_str());
    std::printf(" %s\n", ok ? "-" : err.c_str());
}
int main(int argc, char** argv) {
    show(">seq1 first record\nACGTACGTAC\nGTACGTACGT\nACGT\n>seq2 second\nTTTT\n");
    show(">a\r\nACGT\r\nACGT\r\n>b\r\nTT\r\n");
    show(">empty\n>next\nACGT\n");
    show("");
    std::vector<FastaRecord> r;
    std::string err;
    const bool miss = read_fasta_file(argv[1], &r, &err);
    std::printf("%d %s\n", int(miss), err.c_str());
    const bool got = read_fasta_file(argv[2], &r, &err);
    std::printf("%d %zu %s\n", int(got), r.size(), r.empty() ? "" : r[0].id.c_str());
    return 0;
}
''')
    exe = tmp_path / "fa"
    libdir = os.path.dirname(N.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", str(src), "-I", os.path.join(ROOT, "include"), "-L", libdir,
                           "-lgenoaligner_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)])
    good = tmp_path / "x.fa"
    good.write_bytes(b">x\nACGT\n")
    missing = str(tmp_path / "definitely_not_here_9f3a.fa")
    out = subprocess.check_output([str(exe), missing, str(good)], text=True).splitlines()
    assert out[0] == "1 2 [seq1|first record|ACGTACGTACGTACGTACGTACGT] [seq2|second|TTTT] -"
    assert out[1] == "1 2 [a||ACGTACGT] [b||TT] -"
    assert out[2] == "0 0 record has a header but no sequence (line 1)"
    assert out[3] == "0 0 no records found (line 1)"
    assert out[4] == "0 cannot open '%s'" % missing
    assert out[5] == "1 1 x"



def test_sw_validate_cigar_matches_oracle(oracle):
    """The library's CIGAR check (ga_sw_validate_cigar, the one the batch
    applies) against the oracle's restatement of the reference validator, on
    traced CIGARs and on mutations of them: ops flipped, inserted, dropped or
    replaced by a foreign byte, runs cut, ends and starts moved, the score
    changed, spans past the sequences."""
    import random
    from genoaligner_amd import _native as N
    L = N.lib()
    rng = random.Random(41)
    ok_r, ok_w = C.c_int32(), C.c_int32()
    seen = set()
    for it in range(2500):
        sc = rng.choice([(2, -3, 3, 1), (1, -1, 2, 1), (2, -5, 2, 1), (3, -6, 4, 2)])
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 90)))
        a = rng.randint(0, len(t) - 1)
        p = "".join(c if rng.random() > 0.12 else rng.choice("ACGT") for c in t[a:a + rng.randint(1, 70)])
        if rng.random() < 0.3:
            p = p[:len(p) // 2] + "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 6))) + p[len(p) // 2:]
        score, si, sj, ei, ej, cig = oracle.sw_trace(t, p, sc)
        cig = list(cig)
        k = rng.randrange(9)
        if k == 1 and cig:
            cig[rng.randrange(len(cig))] = rng.choice("MXID")
        elif k == 2:
            cig.insert(rng.randint(0, len(cig)), rng.choice("MXID"))
        elif k == 3 and cig:
            del cig[rng.randrange(len(cig))]
        elif k == 4 and cig:
            cig[rng.randrange(len(cig))] = rng.choice("Z=\0m")
        elif k == 5:
            ei, ej = ei + rng.randint(-2, 2), ej + rng.randint(-2, 2)
        elif k == 6:
            si, sj = si + rng.randint(-1, 2), sj + rng.randint(-1, 2)
        elif k == 7:
            score += rng.choice([-1, 1])
        elif k == 8:
            cig += ["M"] * rng.randint(1, 20)            # past both sequences
        cg = "".join(cig)
        if min(si, sj) < 0:
            continue                                     # negative starts are no SW result
        want = oracle.sw_validate(cg, t, si, ei, p, sj, ej, sc, score)
        scc = N.SWScoringC(*sc)
        assert L.ga_sw_validate_cigar(cg.encode("latin-1"), len(cg), t.encode(), len(t), si, ei, p.encode(),
                                      len(p), sj, ej, C.byref(scc), score, C.byref(ok_r), C.byref(ok_w)) == 0
        got = (bool(ok_r.value), bool(ok_w.value))
        assert got == tuple(bool(x) for x in want), (it, k, cg, t, p, si, ei, sj, ej, score)
        seen.add(got)
    assert seen == {(True, True), (False, True), (False, False)}
    assert L.ga_sw_validate_cigar(None, 3, b"A", 1, 0, 0, b"A", 1, 0, 0, C.byref(N.SWScoringC(1, -1, 1, 1)), 1,
                                  C.byref(ok_r), C.byref(ok_w)) == N.GA_INVALID_ARGUMENT
