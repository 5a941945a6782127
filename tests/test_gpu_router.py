"""GPU tests of the learned router: the sketch and MLP kernels against the CPU
oracle, training on exact-aligner labels, and the invariant that routing never
changes a result (routed == plain, pair for pair)."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairset_from(pairs, smax):
    from genoaligner_amd.workloads import PairSet
    pl = np.array([len(p) for p, _ in pairs], np.int32)
    tl = np.array([len(t) for _, t in pairs], np.int32)
    pa = np.frombuffer("".join(p for p, _ in pairs).encode(), np.uint8).copy()
    ta = np.frombuffer("".join(t for _, t in pairs).encode(), np.uint8).copy()
    po = np.zeros(len(pairs), np.int64)
    to = np.zeros(len(pairs), np.int64)
    po[1:] = np.cumsum(pl)[:-1]
    to[1:] = np.cumsum(tl)[:-1]
    return PairSet(pa, po, pl, ta, to, tl, np.asarray(smax, np.int32), {})


def test_sketch_kernel_matches_oracle(gpu, oracle):
    import torch
    from genoaligner_amd import router as R
    from genoaligner_amd.device import WfaDeviceBatch
    rng = random.Random(12)
    pairs = [("", ""), ("ACG", ""), ("", "ACGTA")]
    for _ in range(200):
        # bytes other than A/C/G/T, some sharing their low bits (a c g t Q E @ S)
        p = "".join(rng.choice("ACGTNacgtQE@S") if rng.random() < 0.05 else rng.choice("ACGT")
                    for _ in range(rng.randint(0, 2500)))
        t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 2500)))
        pairs.append((p, t))
    # sequences of 500,000+ windows take the kernel's 32-bit-histogram path
    big = "".join(rng.choice("ACGT") for _ in range(520_000))
    pairs.append((big, big[: 510_000] + "N" + big[510_001:]))
    pairs.append(("ACGT" * 10, big))
    ps = _pairset_from(pairs, [10] * len(pairs))
    b = WfaDeviceBatch(ps, trace=False)
    feat, lb = R.sketch(b)
    torch.cuda.synchronize()
    feat, lb = feat.cpu().numpy(), lb.cpu().numpy()
    for i, (p, t) in enumerate(pairs):
        f, l = oracle.sketch(p, t)
        assert l == int(lb[i]), i
        assert np.array_equal(f, feat[i]), i


def test_mlp_kernels_match_oracle(gpu, oracle):
    """Forward: same accumulation order as the oracle, so equal to 1e-6.
    Backward: tolerance 1e-5 relative to the gradient scale (expf/logf of the
    device and of the host libm may differ in the last bit)."""
    import torch
    from genoaligner_amd import router as R
    rng = np.random.default_rng(8)
    params = (rng.standard_normal(R.PARAMS) * 0.12).astype(np.float32)
    x = rng.random((300, 64)).astype(np.float32)
    y = rng.integers(0, 16, size=300).astype(np.int32)
    mlp = R.Mlp(params)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    lg = mlp.forward(xd).cpu().numpy()
    h1, h2, elg = oracle.mlp_forward(params, x)
    assert np.allclose(mlp.h1[:300].cpu().numpy(), h1, rtol=1e-6, atol=1e-6)
    assert np.allclose(lg, elg, rtol=1e-6, atol=1e-6)
    # the fused inference forward (fused multiply-adds: within 1e-5 of the
    # oracle's separate multiply and add; 300 rows leave a partial tile)
    fl = mlp.logits_only(xd).cpu().numpy()
    assert np.allclose(fl, elg, rtol=1e-5, atol=1e-5)
    loss = mlp.backward(xd, yd).cpu().numpy()
    g, eloss = oracle.mlp_backward(params, x, y)
    assert np.allclose(loss, eloss, rtol=1e-5, atol=1e-6)
    gd = mlp.grads.cpu().numpy()
    assert np.abs(gd - g).max() <= 1e-5 * np.abs(g).max() + 1e-9
    # one SGD step with momentum: v = mu v + g; p -= lr v
    mlp.mom.fill_(0.25)
    mlp.sgd(0.1, 0.9)
    v = 0.9 * np.float32(0.25) + gd
    assert np.allclose(mlp.params.cpu().numpy(), params - np.float32(0.1) * v.astype(np.float32),
                       rtol=1e-6, atol=1e-7)


def test_route_rule(gpu):
    import torch
    from genoaligner_amd import router as R
    z = np.full((3, 16), -10.0, np.float32)
    z[0, 3] = 5.0                     # certain: bucket 3
    z[1, 2] = 1.0; z[1, 9] = 1.0      # split: argmax 2, 0.9-quantile 9
    z[2, :] = 0.0                     # uniform
    mlp = R.Mlp(R.init_params())
    out0 = torch.empty(3, dtype=torch.int32, device="cuda")
    zd = torch.from_numpy(z).cuda()
    import ctypes as C
    from genoaligner_amd import _native as N
    err = C.c_char_p()
    assert N.lib().ga_mlp_route_device(C.c_void_p(zd.data_ptr()), 3, C.c_float(0.0),
                                       C.c_void_p(out0.data_ptr()), None, C.byref(err)) == 0
    assert out0.cpu().tolist() == [3, 2, 0]
    assert N.lib().ga_mlp_route_device(C.c_void_p(zd.data_ptr()), 3, C.c_float(0.9),
                                       C.c_void_p(out0.data_ptr()), None, C.byref(err)) == 0
    assert out0.cpu().tolist() == [3, 9, 14]
    del mlp


def _results(b):
    sc, cg = b.results()
    return [int(x) for x in sc], cg


def test_routed_equals_plain_for_any_network(gpu, oracle):
    """Untrained, trained-looking and adversarial routers: the routed path
    returns exactly the plain path's scores and CIGARs."""
    import torch
    from genoaligner_amd import router as R
    from genoaligner_amd import workloads as W
    from genoaligner_amd.device import WfaDeviceBatch
    ps = W.wfa_routing(n=1500, lo=60, hi=2500, seed=77)
    want = (np.random.default_rng(2).random(ps.n) < 0.7).astype(np.uint8)
    plain = WfaDeviceBatch(ps, trace=True, want=want)
    plain.run()
    psc, pcg = _results(plain)
    for i in range(0, ps.n, 15):
        osc, ocg = oracle.wfa(ps.a_seq(i), ps.b_seq(i), int(ps.smax[i]))
        assert psc[i] == osc and pcg[i].decode() == (ocg if want[i] else ""), i
    rng = np.random.default_rng(1)
    nets = [R.init_params(3), (rng.standard_normal(R.PARAMS) * 0.3).astype(np.float32)]
    low = np.zeros(R.PARAMS, np.float32)
    low[-16] = 50.0                   # always bucket 0: every pair re-run
    nets.append(low)
    for k, params in enumerate(nets):
        for q in (0.0, 0.9):
            rb = WfaDeviceBatch(ps, trace=True, want=want)
            st = R.run_routed(rb, torch.from_numpy(params).cuda(), q=q)
            rsc, rcg = _results(rb)
            assert rsc == psc, (k, q)
            assert rcg == pcg, (k, q)
            assert st.shortcut + st.routed + st.plain == ps.n
            assert st.rerun <= st.routed
    # score-only batches route too
    sb = WfaDeviceBatch(ps, trace=False)
    R.run_routed(sb, torch.from_numpy(nets[0]).cuda(), q=0.9)
    assert [int(x) for x in sb.results()[0]] == psc


def test_library_router_batches_equal_plain(gpu):
    """With a network loaded into the library (ga_router_set), align_batch_ext
    and align_batch route their WFA batches and return exactly what they
    return without one: the shipped network, an untrained one and one that
    sends every pair to bucket 0 (every pair re-run); mixed CIGAR flags and
    pairs of up to 10 kbp at smax up to 2048."""
    import genoaligner_amd as ga
    from genoaligner_amd import router as R
    from genoaligner_amd import workloads as W
    ps = W.wfa_routing(n=1000, lo=60, hi=2500, seed=78)
    pl = W.wfa_mixed(n=60, lo=5000, hi=10000, seed=79)
    rng = np.random.default_rng(5)

    def reqs_of(p, smax_cap):
        return [ga.AlignRequest(pattern=p.a_seq(i), pattern_len=len(p.a_seq(i)), text=p.b_seq(i),
                                text_len=len(p.b_seq(i)), smax=min(int(p.smax[i]), smax_cap),
                                with_cigar=bool(rng.random() < 0.7)) for i in range(p.n)]

    ext = reqs_of(ps, 4000) + reqs_of(pl, 4000)
    short = reqs_of(ps, 300)[:400]
    fields = ("score", "resolved", "cigar", "rescore_ok", "wellformed_ok")

    def run():
        a, b = ga.align_batch_ext(ext), ga.align_batch(short)
        assert a.ok() and b.ok(), (a.error, b.error)
        return [tuple(getattr(r, f) for f in fields) for r in a.results + b.results]

    ga.set_router(None)
    want = run()
    low = np.zeros(R.PARAMS, np.float32)
    low[-16] = 50.0
    try:
        for params in (R.load(), R.init_params(4), low):
            ga.set_router(params, 0.95)
            assert ga.router_enabled()
            assert run() == want
    finally:
        ga.set_router(None)
    assert not ga.router_enabled()


def test_training_learns_distance_buckets(gpu):
    """A short GPU training run on exact-aligner labels lowers the loss and beats
    the majority-bucket predictor on held-out pairs; routing with the trained
    network re-runs few pairs."""
    import torch
    from genoaligner_amd import router as R
    from genoaligner_amd import workloads as W
    tr = W.wfa_routing(n=20000, lo=100, hi=3000, seed=101)
    te = W.wfa_routing(n=4000, lo=100, hi=3000, seed=202)
    f, y, _ = R.label_batch(tr)
    ft, yt, bt = R.label_batch(te)
    mlp = R.Mlp(R.init_params(5))
    hist = mlp.train(f, y, epochs=4, batch=512, lr=0.05)
    assert hist[-1] < 0.7 * hist[0]
    pred = mlp.route(ft, 0.0).cpu().numpy()
    yt = yt.cpu().numpy()
    acc = float((np.abs(pred - yt) <= 1).mean())
    base = float((np.abs(np.bincount(y.cpu().numpy()).argmax() - yt) <= 1).mean())
    assert acc > base + 0.2, (acc, base)
    st = R.run_routed(bt, mlp.params, q=0.9)
    assert st.routed > 0.5 * te.n
    assert st.miss_rate < 0.15, st
