"""Pins of the router's CPU restatement (oracle/ga_router_oracle.c), which has
no reference counterpart: each check uses an independent computation."""
import random

import numpy as np



def _recount(s):
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    h = np.zeros(256, np.int64)
    for i in range(len(s) - 3):
        w = s[i:i + 4]
        if all(c in code for c in w):
            h[code[w[0]] * 64 + code[w[1]] * 16 + code[w[2]] * 4 + code[w[3]]] += 1
    return h


def test_sketch_matches_recount(oracle):
    rng = random.Random(5)
    for _ in range(60):
        m, n = rng.randint(0, 300), rng.randint(0, 300)
        p = "".join(rng.choice("ACGTN") for _ in range(m))
        t = "".join(rng.choice("ACGTa") for _ in range(n))
        f, lb = oracle.sketch(p, t)
        hp, ht = _recount(p), _recount(t)
        d = np.abs(hp - ht)
        l1 = int(d.sum())
        M = max(m, n, 1)
        assert lb == max(abs(m - n), (l1 + 7) // 8)
        exp = np.zeros(64, np.float32)
        exp[0] = np.float32(m) / np.float32(10000)
        exp[1] = np.float32(n) / np.float32(10000)
        exp[2] = np.float32(abs(m - n)) / np.float32(M)
        exp[3] = np.float32(l1) / np.float32(2 * M)
        for j in range(60):
            exp[4 + j] = np.float32(int(d[j::60].sum()) * 30) / np.float32(M)
        assert np.array_equal(f, exp)


def test_lower_bound_never_exceeds_edit_distance(oracle):
    """||Q4(a) - Q4(b)||_1 <= 8 ED and |m - n| <= ED, on random and adversarial pairs."""
    rng = random.Random(9)
    cases = []
    for _ in range(300):
        m = rng.randint(0, 120)
        p = "".join(rng.choice("ACGT") for _ in range(m))
        if rng.random() < 0.5:
            t = list(p)
            for _ in range(rng.randint(0, 12)):
                op = rng.random()
                j = rng.randint(0, len(t))
                if op < 0.4 and j < len(t):
                    t[j] = rng.choice("ACGT")
                elif op < 0.7:
                    t.insert(j, rng.choice("ACGT"))
                elif j < len(t):
                    del t[j]
            t = "".join(t)
        else:
            t = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 120)))
        cases.append((p, t))
    cases += [("A" * 40, "C" * 40), ("ACGT" * 10, "TGCA" * 10), ("AAAAC", "CAAAA")]
    tight = 0
    for p, t in cases:
        _, lb = oracle.sketch(p, t)
        ed = oracle.levenshtein(p, t)
        assert lb <= ed, (p, t, lb, ed)
        tight += lb == ed
    assert tight > 0


def _np_forward(params, x):
    P = params.astype(np.float64)
    w1 = P[:8192].reshape(64, 128); b1 = P[8192:8320]
    w2 = P[8320:24704].reshape(128, 128); b2 = P[24704:24832]
    w3 = P[24832:26880].reshape(128, 16); b3 = P[26880:26896]
    h1 = np.maximum(x.astype(np.float64) @ w1 + b1, 0)
    h2 = np.maximum(h1 @ w2 + b2, 0)
    return h1, h2, h2 @ w3 + b3


def _np_loss(params, x, y):
    z = _np_forward(params, x)[2]
    z = z - z.max(axis=1, keepdims=True)
    return float(np.mean(np.log(np.exp(z).sum(axis=1)) - z[np.arange(len(y)), y]))


def test_mlp_forward_matches_float64(oracle):
    rng = np.random.default_rng(3)
    params = (rng.standard_normal(26896) * 0.1).astype(np.float32)
    x = rng.random((40, 64)).astype(np.float32)
    h1, h2, lg = oracle.mlp_forward(params, x)
    e1, e2, el = _np_forward(params, x)
    assert np.allclose(h1, e1, rtol=1e-5, atol=1e-5)
    assert np.allclose(h2, e2, rtol=1e-5, atol=1e-5)
    assert np.allclose(lg, el, rtol=1e-4, atol=1e-5)


def test_mlp_gradient_matches_finite_differences(oracle):
    """Analytic FP32 gradient of the oracle vs central differences of a float64
    loss; tolerance 2e-3 relative to the gradient's scale."""
    rng = np.random.default_rng(4)
    params = (rng.standard_normal(26896) * 0.15).astype(np.float32)
    x = rng.random((24, 64)).astype(np.float32)
    y = rng.integers(0, 16, size=24).astype(np.int32)
    g, loss = oracle.mlp_backward(params, x, y)
    assert abs(float(loss.mean()) - _np_loss(params, x, y)) < 1e-4
    idx = np.concatenate([rng.choice(8192, 12, replace=False), 8192 + rng.choice(128, 4, replace=False),
                          8320 + rng.choice(16384, 12, replace=False), 24704 + rng.choice(128, 4, replace=False),
                          24832 + rng.choice(2048, 12, replace=False), 26880 + np.arange(16)])
    h = 1e-3
    scale = float(np.abs(g).max())
    for i in idx:
        pp = params.astype(np.float64).copy(); pp[i] += h
        pm = params.astype(np.float64).copy(); pm[i] -= h
        fd = (_np_loss(pp, x, y) - _np_loss(pm, x, y)) / (2 * h)
        assert abs(fd - g[i]) <= 2e-3 * scale + 1e-5, (int(i), fd, float(g[i]))


def test_planted_gradient_mistake_is_caught(oracle):
    """A dropped ReLU mask in the second layer (planted in a copy of the
    oracle's formula, on the CPU) fails the finite-difference check above."""
    rng = np.random.default_rng(6)
    params = (rng.standard_normal(26896) * 0.15).astype(np.float32)
    x = rng.random((16, 64)).astype(np.float32)
    y = rng.integers(0, 16, size=16).astype(np.int32)
    P = params.astype(np.float64)
    w2 = P[8320:24704].reshape(128, 128); w3 = P[24832:26880].reshape(128, 16)
    h1, h2, z = _np_forward(params, x)
    pz = np.exp(z - z.max(axis=1, keepdims=True)); pz /= pz.sum(axis=1, keepdims=True)
    dl = pz.copy(); dl[np.arange(16), y] -= 1; dl /= 16
    d2_wrong = dl @ w3.T                      # mask omitted
    gw2_wrong = h1.T @ d2_wrong
    g, _ = oracle.mlp_backward(params, x, y)
    assert not np.allclose(gw2_wrong.ravel(), g[8320:24704], atol=1e-4)
    d2 = (dl @ w3.T) * (h2 > 0)
    assert np.allclose((h1.T @ d2).ravel(), g[8320:24704], atol=1e-4)
