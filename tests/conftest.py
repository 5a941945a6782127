import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _unesc(s: str) -> str:
    return "" if s == "-" else s


def read_tsv(name: str):
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([_unesc(x) for x in line.rstrip("\n").split("\t")])
    return rows


def golden_wfa():
    """(smax, pattern, text, score, cigar, flat_score)"""
    return [(int(r[0]), r[1], r[2], int(r[3]), r[4], int(r[5])) for r in read_tsv("wfa_trace.tsv")]


def golden_sw_trace():
    """(scheme, text, pattern, score, start_i, start_j, end_i, end_j, cigar)"""
    out = []
    for r in read_tsv("sw_trace.tsv"):
        sc = tuple(int(x) for x in r[0:4])
        out.append((sc, r[4], r[5], int(r[6]), int(r[7]), int(r[8]), int(r[9]), int(r[10]), r[11]))
    return out


def golden_sw_score():
    """(scheme, text, pattern, score, end_i, end_j) at warp width 64"""
    out = []
    for r in read_tsv("sw_score_w64.tsv"):
        sc = tuple(int(x) for x in r[0:4])
        out.append((sc, r[4], r[5], int(r[6]), int(r[7]), int(r[8])))
    return out


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible GPU; skips when no GPU is present."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from genoaligner_amd import _native
    _native.lib()           # raises if the library was not built
    return torch.device("cuda:0")
