"""The request loop on the GPU: every request's results in the concurrent mix
equal the results of the same request run alone (integer results: exact)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    fields = ("score", "resolved", "cigar", "rescore_ok", "wellformed_ok")
    extra = ("start_i", "start_j", "end_i", "end_j", "too_large")
    for x, y in zip(a, b):
        for f in fields + (extra if hasattr(x, "end_i") else ()):
            if getattr(x, f) != getattr(y, f):
                return False
    return len(a) == len(b)


@pytest.mark.parametrize("route", [False, True])
def test_results_alone_equal_results_in_the_mix(gpu, oracle, route):
    """route=True: the slots' WFA batches take the library's routed path (the
    shipped network loaded); the alone results stay unrouted."""
    import genoaligner_amd as ga
    from genoaligner_amd import workloads as W
    from genoaligner_amd.serve import LibraryBackend, Server
    rng = np.random.default_rng(31)
    reqs = [W.serving_request(rng, ga, max_pairs=48, wfa_len=(100, 1500), sw_q=(50, 300), sw_t=(50, 500))
            for _ in range(30)]
    m = ga.blosum62(11, 1)
    try:
        _mix_vs_alone(ga, W, Server(budget_bytes=4 << 30, slots=3, start=False,
                                    backend=LibraryBackend(route=route)), reqs, m, rng, oracle)
    finally:
        ga.set_router(None)


def _mix_vs_alone(ga, W, s, reqs, m, rng, oracle):
    tickets = [s.submit(kind, items, with_cigar=cig) for kind, items, cig in reqs[:12]]
    s.start()                                         # 12 queued requests batch together
    for kind, items, cig in reqs[12:]:                # the rest arrive while slots are busy
        tickets.append(s.submit(kind, items, with_cigar=cig))
        time.sleep(float(rng.exponential(0.002)))
    for t in tickets:
        assert t.wait(120)
    s.close()
    ga.set_router(None)                               # the alone runs take the plain path
    assert len(s.batches) < len(reqs)                 # requests were batched together
    for (kind, items, cig), t in zip(reqs, tickets):
        assert t.error is None, t.error
        alone = ga.align_batch_ext(items) if kind == "wfa" else ga.align_sw_matrix_batch(items, m)
        assert alone.ok()
        assert _same(t.results, alone.results), kind
    # and the alone results are the oracle's, on a sample
    kind, items, cig = next(r for r in reqs if r[0] == "wfa")
    for it, res in list(zip(items, tickets[reqs.index((kind, items, cig))].results))[:10]:
        sc, cg = oracle.wfa(it.pattern, it.text, it.smax)
        assert (res.score, res.cigar) == (sc, cg if cig else "")
