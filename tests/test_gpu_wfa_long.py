"""The top WFA size class (smax above 1,024: per-workgroup stacks of up to
(4,001)^2 offsets, 4 or 8 waves per workgroup) on pairs of 5-10 kbp: the
mixed-batch generator's long end, pairs near and past their bound, a few at
bounds up to the extension's 4,000 and some whose bound is their exact
distance, 70 % with CIGAR.  The same batch runs through the host entry point,
the device entry point with host-planned and with device-planned size
classes, and the routed path (trained and untrained networks); every score
and CIGAR equals the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def long_batch(oracle):
    from genoaligner_amd import workloads as W
    rng = np.random.default_rng(2027)
    n = 104
    div = rng.uniform(0.0, 0.15, size=n)
    div[84:92] = rng.uniform(0.2, 0.3, size=8)        # near or past min(0.2 len, 2048)
    div[92:96] = rng.uniform(0.35, 0.45, size=4)      # distances ~2,000-3,500
    ps = W.wfa_mixed(n=n, lo=5200, hi=10_000, seed=2027, div=div)
    ps.smax[92:96] = [2500, 3000, 3500, 4000]
    # bound = exact distance (the last one short by one: unresolved)
    for i, delta in zip(range(96, 104), (0, 0, 0, 0, 0, 0, -1, -1)):
        ps.smax[i] = min(4000, max(0, oracle.levenshtein(ps.a_seq(i), ps.b_seq(i)) + delta))
    assert (ps.smax[:92] > 1024).all()
    want = (rng.random(n) < 0.7).astype(np.uint8)
    exp = []
    for i in range(n):
        sc, cg = oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i]))
        exp.append((sc, cg if want[i] else ""))
    scores = np.array([e[0] for e in exp])
    assert (scores >= 0).sum() >= 90 and (scores < 0).sum() >= 3
    assert scores.max() > 2048
    return ps, want, exp


def _dev_results(b):
    sc, cg = b.results()
    return [(int(sc[i]), cg[i].decode()) for i in range(b.n)]


def test_long_pairs_host_entry(gpu, long_batch):
    import genoaligner_amd as ga
    ps, want, exp = long_batch
    reqs = [ga.AlignRequest(pattern=ps.a_seq(i).decode(), pattern_len=int(ps.a_len[i]),
                            text=ps.b_seq(i).decode(), text_len=int(ps.b_len[i]), smax=int(ps.smax[i]),
                            with_cigar=bool(want[i])) for i in range(ps.n)]
    b = ga.align_batch_ext(reqs)
    assert b.ok(), b.error
    assert [(r.score, r.cigar) for r in b.results] == exp
    assert all(r.rescore_ok and r.wellformed_ok for r in b.results if r.cigar)


@pytest.mark.parametrize("classes", [True, False])
def test_long_pairs_device_entry(gpu, long_batch, classes):
    from genoaligner_amd.device import WfaDeviceBatch
    ps, want, exp = long_batch
    b = WfaDeviceBatch(ps, trace=True, want=want, classes=classes)
    b.run()
    assert _dev_results(b) == exp
    b.run()                                   # workspace reuse: same results again
    assert _dev_results(b) == exp


@pytest.mark.parametrize("trained", [True, False])
def test_long_pairs_routed(gpu, long_batch, trained):
    import torch
    from genoaligner_amd import router as R
    from genoaligner_amd.device import WfaDeviceBatch
    ps, want, exp = long_batch
    params = torch.from_numpy(R.load() if trained else R.init_params(seed=3)).cuda()
    b = WfaDeviceBatch(ps, trace=True, want=want)
    st = R.run_routed(b, params, q=0.95)
    assert _dev_results(b) == exp
    assert st.shortcut + st.routed + st.plain == ps.n
