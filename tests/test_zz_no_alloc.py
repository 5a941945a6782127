"""ga_wfa_device_noalloc (the WFA device entry point whose scratch lives in a
static device array, for callers that cannot allocate) against
ga_wfa_device on the same batches: every score and CIGAR equal, on calls in
a row whose inputs shrink (the array keeps the last call's contents), with and
without CIGARs, plus its size refusals."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixed(n, seed, lo=100, hi=3000):
    from genoaligner_amd import workloads as W
    ps = W.wfa_mixed(n=n, lo=lo, hi=hi, seed=seed)
    return ps


def _long(n, seed):
    from genoaligner_amd import workloads as W
    return W.wfa_mixed(n=n, lo=5200, hi=8000, seed=seed)


def _concat(a, b):
    from genoaligner_amd.workloads import PairSet
    def cat(x, xo, xl, y, yo, yl):
        return np.concatenate([x, y]), np.concatenate([xo, yo + len(x)]), np.concatenate([xl, yl])
    pa = cat(a.a, a.a_off, a.a_len, b.a, b.a_off, b.a_len)
    pb = cat(a.b, a.b_off, a.b_len, b.b, b.b_off, b.b_len)
    return PairSet(*pa, *pb, np.concatenate([a.smax, b.smax]).astype(np.int32))


def _noalloc(batch):
    from genoaligner_amd import _native as N
    err = C.c_char_p()
    rc = N.lib().ga_wfa_device_noalloc(C.byref(batch.b), None, C.byref(err))
    assert rc == N.GA_OK, err.value
    return batch.results()


def _pair(ps, want, trace=True):
    from genoaligner_amd.device import WfaDeviceBatch
    ref = WfaDeviceBatch(ps, trace=trace, want=want if trace else None)
    ref.run()
    got = WfaDeviceBatch(ps, trace=trace, want=want if trace else None, classes=False)
    return ref.results(), _noalloc(got)


def _same(a, b):
    (sa, ca), (sb, cb) = a, b
    assert np.array_equal(sa, sb)
    if ca is not None:
        assert ca == cb


def test_noalloc_matches_device_entry_on_shrinking_calls(gpu):
    big = _concat(_mixed(3000, seed=41), _long(24, seed=42))
    want = (np.random.default_rng(1).random(big.n) < 0.7).astype(np.uint8)
    ref, got = _pair(big, want)
    assert (ref[0] >= 0).sum() > 2900
    _same(ref, got)
    small = _mixed(300, seed=43, lo=100, hi=1000)
    ref, got = _pair(small, np.ones(small.n, dtype=np.uint8))
    _same(ref, got)
    # score only, then CIGAR again on another small batch
    ref, got = _pair(_mixed(500, seed=44), None, trace=False)
    _same(ref, got)
    tiny = _mixed(7, seed=45, lo=100, hi=400)
    ref, got = _pair(tiny, np.ones(tiny.n, dtype=np.uint8))
    _same(ref, got)


def test_noalloc_refusals(gpu):
    from genoaligner_amd import _native as N
    L = N.lib()
    err = C.c_char_p()
    b = N.WfaDevBatch()
    b.n_pairs = (1 << 20) + 1
    b.smax_max = 64
    b.max_len = 1000
    assert L.ga_wfa_device_noalloc(C.byref(b), None, C.byref(err)) == N.GA_INVALID_ARGUMENT
    assert b"GA_WFA_NOALLOC_MAX_PAIRS" in err.value
    b = N.WfaDevBatch()
    b.n_pairs = 1
    b.smax_max = 4000
    b.max_len = 40000
    b.cigar = C.c_void_p(1)               # refused on its size before any pointer is used
    assert L.ga_wfa_device_noalloc(C.byref(b), None, C.byref(err)) == N.GA_INVALID_ARGUMENT
    assert b"scratch" in err.value
    b.n_pairs = 0
    b.smax_max = 4001
    assert L.ga_wfa_device_noalloc(C.byref(b), None, C.byref(err)) == N.GA_INVALID_ARGUMENT
