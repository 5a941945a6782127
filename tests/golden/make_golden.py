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


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libga_ref.so missing: run oracle/build_ref.sh first")
    for fn in (write_wfa, write_sw_trace, write_sw_score):
        print("wrote", fn())


if __name__ == "__main__":
    main()
