#!/usr/bin/env python3
"""Write the golden vectors under tests/golden/ from the reference's own code.

The reference's kernels and API are compiled for the CPU through the
reference's HIP CPU shim by oracle/build_ref.sh into oracle/_ref/libga_ref.so;
this script drives that library on seeded inputs and writes inputs + outputs as
TSV.  The fixtures are data only (inputs and the reference's outputs); the
script is re-runnable wherever the reference tree is present:

    oracle/build_ref.sh && python3 tests/golden/make_golden.py

Case families follow the reference's own tests: the hand cases of
tests/parity/traceback_cpu.cpp:344-359 and tests/parity/r1_check.cpp:117-131,
the API known answers of tests/api/test_api.cpp:104-119, :196-218, :296-330,
:402-477, and seeded random families in the style of tests/parity/wfa_parity.cpp
(seed 12345) and tests/sw/test_sw_parity.cpp (seed 20260912).
"""
from __future__ import annotations

import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

ALPHA = "ACGT"


def mutate(rng, s, rate, alpha=ALPHA):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 2:
            out.append(rng.choice([a for a in alpha if a != c] or [c]))
        elif r < rate * 3 / 4:
            out.append(c)
            out.append(rng.choice(alpha))
        elif r < rate:
            pass
        else:
            out.append(c)
    return "".join(out)


def rand_seq(rng, n, alpha=ALPHA):
    return "".join(rng.choice(alpha) for _ in range(n))


def esc(s: str) -> str:
    return s if s else "-"


def wfa_cases():
    hand = [
        ("ACGT", "ACGT"), ("ACGT", "ACGA"), ("ACGTA", "ACGT"), ("ACGT", "ACGTA"),
        ("GG", "TTT"), ("C", "GT"), ("CT", "CCC"), ("CG", "GAC"), ("A", "AA"),
        ("", ""), ("ACGT", ""), ("", "ACGT"), ("AAAA", "AAAAA"), ("ACGTACGT", "TTTTTTTT"),
        ("A", "A"), ("A", "T"), ("AA", "A"), ("ACGTACGT", "ACGTTCGT"),
        # literal byte semantics: N matches N only, case is significant
        ("ACGNNACGT", "ACGNNACGT"), ("acgt", "ACGT"), ("ACNT", "ACGT"),
    ]
    cases = [(p, t, s) for p, t in hand for s in (0, 1, 3, 8, 64)]
    # heterogeneous-length single-substitution family (test_api.cpp:296-330)
    for L in (8, 40, 120, 300, 55):
        t = "".join("ACGT"[i % 4] for i in range(L))
        p = list(t)
        p[L // 2] = "C" if p[L // 2] == "A" else "A"
        cases.append(("".join(p), t, 64))
    rng = random.Random(12345)
    for i in range(400):
        L = rng.choice([rng.randint(1, 12), rng.randint(20, 200), rng.randint(200, 600)])
        p = rand_seq(rng, L)
        rate = rng.choice([0.0, 0.01, 0.02, 0.05, 0.1, 0.2])
        t = mutate(rng, p, rate)
        if rng.random() < 0.5:
            p, t = t, p
        smax = rng.choice([2, 8, 16, 32, 64, 128, 200, 511])
        cases.append((p, t, smax))
    # exhaustive-style tiny family (r1_check.cpp:164-180 style, seed 2024)
    rng = random.Random(2024)
    for i in range(300):
        p = rand_seq(rng, rng.randint(0, 6), "AC")
        t = rand_seq(rng, rng.randint(0, 6), "AC")
        cases.append((p, t, 6))
    return cases


def write_wfa():
    path = os.path.join(HERE, "wfa_trace.tsv")
    with open(path, "w") as f:
        f.write("# reference wfa_trace_kernel (score, CIGAR) via its CPU shim\n")
        f.write("#smax\tpattern\ttext\tscore\tcigar\tflat_score\n")
        for p, t, s in wfa_cases():
            sc, cg = O.ref_wfa(p, t, s)
            fl = O.ref_wfa_score_flat(p, t, s)
            f.write(f"{s}\t{esc(p)}\t{esc(t)}\t{sc}\t{esc(cg)}\t{fl}\n")
    return path


SCHEMES = [(2, -3, 5, 2), (1, -1, 2, 1), (2, -3, 3, 1), (3, -1, 2, 2), (5, -4, 10, 1)]


def sw_cases():
    cases = [
        ("TTTACGTGTT", "ACGTGT", (2, -3, 5, 2)),
        ("ACGTACGT", "ACGTACGT", (2, -3, 5, 2)),
        ("AACCGGTT" + "C" * 6 + "TTGGCCAA", "AACCGGTTTTGGCCAA", (2, -3, 5, 2)),
        ("AAAAAAAA", "CCCCCCCC", (2, -3, 5, 2)),
        ("ACGT", "", (2, -3, 5, 2)), ("", "ACGT", (2, -3, 5, 2)), ("", "", (2, -3, 5, 2)),
        ("ACGT", "ACGT", (1, -1, 2, 1)), ("ACNNGT", "ACNNGT", (2, -3, 3, 1)),
        ("acgtACGT", "ACGTACGT", (2, -3, 3, 1)),
    ]
    rng = random.Random(20260912)
    for i in range(300):
        sc = rng.choice(SCHEMES)
        kind = rng.random()
        if kind < 0.5:
            t = rand_seq(rng, rng.randint(1, 200))
            a = rng.randint(0, len(t))
            b = rng.randint(a, min(len(t), a + rng.randint(1, 120)))
            p = mutate(rng, t[a:b], rng.choice([0.0, 0.05, 0.1, 0.2]))
            if not p:
                p = rand_seq(rng, 5)
        elif kind < 0.8:
            t = rand_seq(rng, rng.randint(1, 150))
            p = rand_seq(rng, rng.randint(1, 150))
        else:
            t = rand_seq(rng, rng.randint(1, 40), "AC")
            p = rand_seq(rng, rng.randint(1, 40), "AC")
        cases.append((t, p, sc))
    # an embedded mutated copy, the 256 x 1024 workload's shape at smaller size
    rng = random.Random(20261129)
    for i in range(40):
        q = rand_seq(rng, rng.randint(64, 256))
        t = rand_seq(rng, rng.randint(300, 600))
        off = rng.randint(0, len(t))
        t = t[:off] + mutate(rng, q, 0.1) + t[off:]
        cases.append((t, q, (2, -3, 3, 1)))
    return cases


def write_sw_trace():
    path = os.path.join(HERE, "sw_trace.tsv")
    with open(path, "w") as f:
        f.write("# reference sw_trace_kernel + sw_trace_walk via its CPU shim\n")
        f.write("#match\tmismatch\tgap_open\tgap_extend\ttext\tpattern\tscore\tstart_i\tstart_j\tend_i\tend_j\tcigar\n")
        for t, p, sc in sw_cases():
            r = O.ref_sw_trace(t, p, sc)
            f.write("\t".join(str(x) for x in sc) + f"\t{esc(t)}\t{esc(p)}\t"
                    + "\t".join(str(x) for x in r[:5]) + f"\t{esc(r[5])}\n")
    return path


def write_sw_score():
    """sw_score_kernel at warpSize 64 (the AMD width) -- its end coordinate on
    ties follows the lane rule, which only the kernel itself pins."""
    rng = random.Random(7)
    cases = []
    for i in range(60):
        sc = rng.choice(SCHEMES[:3] + [(2, -3, 1, 2)])          # includes ge > go (score-only)
        if i % 3 == 0:   # low-complexity: many ties
            t = rand_seq(rng, rng.randint(1, 160), "AC")
            p = rand_seq(rng, rng.randint(1, 150), "A")
        else:
            t = rand_seq(rng, rng.randint(0, 200))
            p = rand_seq(rng, rng.randint(0, 170))
        cases.append((t, p, sc))
    path = os.path.join(HERE, "sw_score_w64.tsv")
    with open(path, "w") as f:
        f.write("# reference sw_score_kernel under warp emulation, warpSize 64\n")
        f.write("#match\tmismatch\tgap_open\tgap_extend\ttext\tpattern\tscore\tend_i\tend_j\n")
        by_scheme = {}
        for k, (t, p, sc) in enumerate(cases):
            by_scheme.setdefault(sc, []).append(k)
        res = [None] * len(cases)
        for sc, ks in by_scheme.items():
            rr = O.ref_sw_score([(cases[k][0], cases[k][1]) for k in ks], sc, 64)
            for k, r in zip(ks, rr):
                res[k] = r
        for (t, p, sc), r in zip(cases, res):
            f.write("\t".join(str(x) for x in sc) + f"\t{esc(t)}\t{esc(p)}\t"
                    + "\t".join(str(x) for x in r) + "\n")
    return path


# ---------------------------------------------------------------- MSA
MSA_PARAMS = {
    "default": O.MsaParams(),
    "legacy": O.MsaParams(psgp=False),
    "global": O.MsaParams(free_end_gaps=False),
    "protein": O.MsaParams.protein(),
    "protein_semi": O.MsaParams(gap_open=11.0, gap_extend=1.0, free_end_gaps=True, alpha=20,
                                sub=O.MSA_BLOSUM62.copy()),
}
AA = "ARNDCQEGHILKMFPSTWYV"


def msa_mutate(rng, s, ps, pi, pd, alpha=ALPHA):
    """substitutions + small indels (percent), in the style of
    tests/parity/msa_pp_parity.cpp:141-153"""
    out = []
    for c in s:
        r = rng.randrange(100)
        if r < pd:
            continue
        out.append(rng.choice(alpha) if r < pd + ps else c)
        if rng.randrange(100) < pi:
            out.append(rng.choice(alpha))
    return "".join(out)


def msa_merged(rng, tag, rows_list):
    """Progressively merge single-row profiles with the reference's alignment
    (rows of merge_profiles), giving fractional columns and gap columns."""
    rows = [rows_list[0]]
    for r in rows_list[1:]:
        sc, span, cg = O.ref_msa_align(rows, [r], MSA_PARAMS[tag])
        rows = O.msa_merge_rows(rows, [r], span, cg)
    return rows


def msa_cases():
    """(params tag, rows A, rows B): the case families of the reference's
    tests/parity/msa_pp_parity.cpp:155-237 (hand pairs, merged profiles,
    global mode, random sweeps of merged and single-sequence profiles,
    protein), plus longer profiles (several 64-column strips) and many-row
    profiles."""
    rng = random.Random(4242)
    out = []
    hand = [("ACGTACGT", "ACGTACGT"), ("ACGTACGT", "ACGTTTTACGT"),
            ("ACGTACGTACGTACGTACGT", "GTACGTAC"), ("ANRYSWKB", "ACGTACGT"), ("A", "T"),
            ("ACGT", "A"), ("acgtNNNN", "ACGTACGT"), ("ACGTACGT", "TTTT")]
    for tag in ("default", "legacy", "global"):
        for a, b in hand:
            out.append((tag, [a], [b]))
        ab = msa_merged(rng, tag, ["ACGTACGT", "ACGTTCGT"])
        out.append((tag, ab, ["ACGTACGTAA"]))
        out.append((tag, ab, ["TTGCGCGC"]))
        for t in range(40):
            anc = rand_seq(rng, 20 + rng.randrange(60))
            A = msa_merged(rng, tag, [msa_mutate(rng, anc, 10, 5, 5), msa_mutate(rng, anc, 10, 5, 5)])
            anc2 = rand_seq(rng, 20 + rng.randrange(60))
            B = msa_merged(rng, tag, [msa_mutate(rng, anc2, 10, 5, 5), msa_mutate(rng, anc2, 10, 5, 5)])
            out.append((tag, A, B))
        for t in range(30):
            out.append((tag, [rand_seq(rng, 5 + rng.randrange(60))], [rand_seq(rng, 5 + rng.randrange(60))]))
        for t in range(12):                                   # several strips, related
            anc = rand_seq(rng, 100 + rng.randrange(400))
            k = 1 + rng.randrange(5)
            A = msa_merged(rng, tag, [msa_mutate(rng, anc, 8, 3, 3) for _ in range(k)])
            B = msa_merged(rng, tag, [msa_mutate(rng, anc, 8, 3, 3) for _ in range(1 + rng.randrange(4))])
            if rng.random() < 0.3:                            # a fragment
                lo = rng.randrange(len(B[0]) // 2)
                B = [r[lo: lo + len(B[0]) // 3] for r in B]
            out.append((tag, A, B))
    for tag in ("protein", "protein_semi"):
        out.append((tag, ["MVLSPADKTNVKAAWGKVGAHAGEYG"], ["VHWTAEEKQLITGLWGKVNVAECGA"]))
        out.append((tag, ["AXBZJUO"], ["ACDEFGH"]))
        for t in range(40):
            out.append((tag, [rand_seq(rng, 5 + rng.randrange(60), AA)],
                        [rand_seq(rng, 5 + rng.randrange(60), AA)]))
        a = msa_merged(rng, tag, [rand_seq(rng, 30, AA), rand_seq(rng, 30, AA)])
        out.append((tag, a, [rand_seq(rng, 35, AA)]))
        for t in range(8):
            anc = rand_seq(rng, 80 + rng.randrange(300), AA)
            A = msa_merged(rng, tag, [msa_mutate(rng, anc, 20, 3, 3, AA) for _ in range(1 + rng.randrange(4))])
            B = msa_merged(rng, tag, [msa_mutate(rng, anc, 20, 3, 3, AA) for _ in range(1 + rng.randrange(3))])
            out.append((tag, A, B))
    return out


def f32hex(x) -> str:
    import struct
    return "%08x" % struct.unpack("<I", struct.pack("<f", float(x)))[0]


def write_msa_pp():
    path = os.path.join(HERE, "msa_pp.tsv")
    with open(path, "w") as f:
        f.write("# reference genomsa::align_profiles (src/msa/msa_ref.cpp) via its CPU shim; the shipped\n"
                "# msa_pp_trace_kernel body agrees on every row (checked when written)\n"
                "#params\trows_a\trows_b\tscore_f32hex\tai\taj\tbi\tbj\tcigar\n")
        for tag, A, B in msa_cases():
            P = MSA_PARAMS[tag]
            sc, span, cg = O.ref_msa_align(A, B, P)
            assert O.ref_msa_align(A, B, P, kernel=True) == (sc, span, cg), (tag, A, B)
            f.write("\t".join([tag, ",".join(A), ",".join(B), f32hex(sc), *map(str, span), esc(cg)]) + "\n")
    return path


def kmer_nj_cases():
    """(k, alpha, sequences): related families (so NJ has structure), random
    sets, IUPAC / ambiguous letters that break k-mers, duplicate sequences
    (distance 0, NJ ties) and short sequences with no k-mer at all."""
    rng = random.Random(20261129)
    out = []
    for t in range(10):
        anc = rand_seq(rng, 200 + rng.randrange(300))
        fam = [msa_mutate(rng, anc, 8, 2, 2) for _ in range(4 + rng.randrange(20))]
        out.append((5, 4, fam))
    for t in range(6):
        n = 3 + rng.randrange(30)
        out.append((rng.choice([3, 4, 5, 6, 8]), 4, [rand_seq(rng, 10 + rng.randrange(120)) for _ in range(n)]))
    seqs = [rand_seq(rng, 60) for _ in range(5)]
    seqs += [seqs[0], seqs[2], "ACGNNACGTRYACGT", "AC", "", "acgtacgtacgt", "ACGTACGTACGT"]
    out.append((5, 4, seqs))
    out.append((3, 4, seqs))
    for t in range(6):
        anc = rand_seq(rng, 100 + rng.randrange(200), AA)
        fam = [msa_mutate(rng, anc, 15, 2, 2, AA) for _ in range(4 + rng.randrange(16))]
        fam.append("BZJXUO" + fam[0][:20])
        out.append((2, 20, fam))
    return out


def write_kmer_nj():
    path = os.path.join(HERE, "kmer_nj.tsv")
    with open(path, "w") as f:
        f.write("# reference genomsa::kmer_distances and genomsa::nj_tree (src/msa/msa_ref.cpp) via its CPU shim\n"
                "#k\talpha\tsequences\tdistances_f32hex\tchildren (internal nodes n..2n-2)\n")
        for k, alpha, seqs in kmer_nj_cases():
            D = O.ref_kmer_distances(seqs, k, alpha)
            ch = O.ref_nj_tree(D, len(seqs)) if len(seqs) >= 2 else []
            f.write("\t".join([str(k), str(alpha), ",".join(esc(s) for s in seqs),
                               ",".join(f32hex(x) for x in D),
                               ",".join(f"{a}:{b}" for a, b in ch)]) + "\n")
    return path


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libga_ref.so missing: run oracle/build_ref.sh first")
    for fn in (write_wfa, write_sw_trace, write_sw_score, write_msa_pp, write_kmer_nj):
        print("wrote", fn())


if __name__ == "__main__":
    main()
