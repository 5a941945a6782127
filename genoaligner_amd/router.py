"""Learned router: a 4-mer/length sketch per pair and a small FP32 MLP
(64 -> 128 -> 128 -> 16) that predicts the pair's edit-distance bucket, trained
on the GPU by the library's own forward/backward/SGD kernels on labels the
exact aligner produces.

The router never changes a result.  It only picks the bound (and so the kernel
size class) a pair is first aligned under; the library re-runs every pair that
is not resolved under a routed bound with the pair's own smax, and marks a pair
unresolved without aligning it only when the exact q-gram lower bound
max(|m - n|, ceil(||Q4(P) - Q4(T)||_1 / 8)) already exceeds its smax.

Device memory is torch tensors used as plain buffers; every computation is a
kernel of libgenoaligner_amd (entry points ga_sketch_device, ga_mlp_*_device,
ga_wfa_routed_device).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _native as N
from .device import WfaDeviceBatch, _check, _stream_handle
from .workloads import PairSet

SKETCH_DIM, HIDDEN, BUCKETS, PARAMS = 64, 128, 16, 26896
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "router_wfa_routing.npy")


def bucket_bounds() -> np.ndarray:
    lib = N.lib()
    return np.array([lib.ga_route_bucket_bound(b) for b in range(BUCKETS)], dtype=np.int64)


def labels_from_scores(scores: np.ndarray) -> np.ndarray:
    """Bucket of each exact result: the smallest bucket whose bound holds the
    distance; unresolved (-1) -> the last bucket."""
    bounds = bucket_bounds()
    sc = np.asarray(scores, dtype=np.int64)
    lab = np.searchsorted(bounds[:-1], np.maximum(sc, 0), side="left").astype(np.int32)
    lab[sc < 0] = BUCKETS - 1
    return lab


def init_params(seed: int = 1) -> np.ndarray:
    """He-normal weights, zero biases, in the library's packed layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((SKETCH_DIM, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, BUCKETS)):
        parts.append((rng.standard_normal((fan_in, fan_out)) * np.sqrt(2.0 / fan_in)).astype(np.float32).ravel())
        parts.append(np.zeros(fan_out, np.float32))
    p = np.concatenate(parts)
    assert p.size == PARAMS
    return p


def _p(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def sketch(batch: WfaDeviceBatch, stream=None):
    """(features [n, 64] float32, lower bound [n] int32) on the device."""
    dev = batch.score.device
    feat = torch.empty((batch.n, SKETCH_DIM), dtype=torch.float32, device=dev)
    lb = torch.empty(batch.n, dtype=torch.int32, device=dev)
    err = C.c_char_p()
    _check(N.lib().ga_sketch_device(C.byref(batch.b), _p(feat), _p(lb), C.c_void_p(_stream_handle(stream)),
                                    C.byref(err)), err, "ga_sketch_device")
    return feat, lb


class Mlp:
    """Parameters and scratch of the MLP on one device."""

    def __init__(self, params: np.ndarray, device: str = "cuda:0"):
        self.dev = torch.device(device)
        self.params = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(self.dev)
        self.grads = torch.zeros(PARAMS, dtype=torch.float32, device=self.dev)
        self.mom = torch.zeros(PARAMS, dtype=torch.float32, device=self.dev)
        self._cap = 0

    def _scratch(self, n: int):
        if n > self._cap:
            self._cap = n
            self.h1 = torch.empty((n, HIDDEN), dtype=torch.float32, device=self.dev)
            self.h2 = torch.empty((n, HIDDEN), dtype=torch.float32, device=self.dev)
            self.logits = torch.empty((n, BUCKETS), dtype=torch.float32, device=self.dev)
            self.loss = torch.empty(n, dtype=torch.float32, device=self.dev)
            self.work = torch.empty(n * (BUCKETS + 2 * HIDDEN), dtype=torch.float32, device=self.dev)

    def forward(self, feat: torch.Tensor, stream=None) -> torch.Tensor:
        n = feat.shape[0]
        self._scratch(n)
        err = C.c_char_p()
        _check(N.lib().ga_mlp_forward_device(_p(self.params), _p(feat), n, _p(self.h1), _p(self.h2),
                                             _p(self.logits), C.c_void_p(_stream_handle(stream)), C.byref(err)),
               err, "ga_mlp_forward_device")
        return self.logits[:n]

    def logits_only(self, feat: torch.Tensor, stream=None) -> torch.Tensor:
        """Inference: the logits by the fused kernel (no activations kept)."""
        n = feat.shape[0]
        out = torch.empty((n, BUCKETS), dtype=torch.float32, device=self.dev)
        err = C.c_char_p()
        _check(N.lib().ga_mlp_logits_device(_p(self.params), _p(feat), n, _p(out),
                                            C.c_void_p(_stream_handle(stream)), C.byref(err)),
               err, "ga_mlp_logits_device")
        return out

    def backward(self, feat: torch.Tensor, label: torch.Tensor, stream=None) -> torch.Tensor:
        """Gradients of the mean cross-entropy of the last forward() (same
        rows) into self.grads; returns the per-sample loss."""
        n = feat.shape[0]
        err = C.c_char_p()
        _check(N.lib().ga_mlp_backward_device(_p(self.params), _p(feat), _p(self.h1), _p(self.h2),
                                              _p(self.logits), _p(label), n, _p(self.grads), _p(self.loss),
                                              _p(self.work), C.c_void_p(_stream_handle(stream)), C.byref(err)),
               err, "ga_mlp_backward_device")
        return self.loss[:n]

    def sgd(self, lr: float, mu: float, stream=None):
        err = C.c_char_p()
        _check(N.lib().ga_mlp_sgd_device(_p(self.params), _p(self.grads), _p(self.mom), C.c_float(lr),
                                         C.c_float(mu), C.c_void_p(_stream_handle(stream)), C.byref(err)),
               err, "ga_mlp_sgd_device")

    def route(self, feat: torch.Tensor, q: float) -> torch.Tensor:
        logits = self.forward(feat)
        n = feat.shape[0]
        out = torch.empty(n, dtype=torch.int32, device=self.dev)
        err = C.c_char_p()
        _check(N.lib().ga_mlp_route_device(_p(logits), n, C.c_float(q), _p(out),
                                           C.c_void_p(_stream_handle(None)), C.byref(err)),
               err, "ga_mlp_route_device")
        return out

    def train(self, feat: torch.Tensor, label: torch.Tensor, epochs: int = 4, batch: int = 1024,
              lr: float = 0.05, mu: float = 0.9, seed: int = 7, log=None) -> List[float]:
        """Minibatch SGD with momentum; returns the mean loss of each epoch."""
        n = feat.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            total = torch.zeros((), dtype=torch.float64, device=self.dev)
            for a in range(0, n - batch + 1, batch):
                idx = perm[a:a + batch]
                xb = feat.index_select(0, idx).contiguous()
                yb = label.index_select(0, idx).contiguous()
                self.forward(xb)
                loss = self.backward(xb, yb)
                self.sgd(lr, mu)
                total += loss.sum(dtype=torch.float64)
            hist.append(float(total.item()) / max(1, (n // batch) * batch))
            if log:
                log(f"epoch {ep}: mean loss {hist[-1]:.4f}")
        return hist

    def numpy(self) -> np.ndarray:
        return self.params.cpu().numpy().copy()


def label_batch(ps: PairSet, device: str = "cuda:0"):
    """Features and exact-aligner labels of a PairSet, on the device: the
    aligner runs score-only with each pair's own smax."""
    b = WfaDeviceBatch(ps, trace=False, device=device)
    b.run()
    feat, _ = sketch(b)
    torch.cuda.synchronize()
    labels = torch.from_numpy(labels_from_scores(b.score.cpu().numpy())).to(feat.device)
    return feat, labels, b


@dataclass
class RouteStats:
    shortcut: int
    routed: int
    plain: int
    rerun: int

    @property
    def miss_rate(self) -> float:
        return self.rerun / self.routed if self.routed else 0.0


def run_routed(batch: WfaDeviceBatch, params: torch.Tensor, q: float = 0.9, stream=None,
               stats: bool = True) -> Optional[RouteStats]:
    """ga_wfa_routed_device over a device batch (results in batch.score /
    batch.cigar as for batch.run()).  With stats the call waits for the stream
    to report them; without, it only enqueues and returns None."""
    st = N.RouteStats()
    err = C.c_char_p()
    _check(N.lib().ga_wfa_routed_device(C.byref(batch.b), _p(params), C.c_float(q),
                                        C.byref(st) if stats else None,
                                        C.c_void_p(_stream_handle(stream)), C.byref(err)),
           err, "ga_wfa_routed_device")
    return RouteStats(st.shortcut, st.routed, st.plain, st.rerun) if stats else None


def save(params: np.ndarray, path: str = WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.ascontiguousarray(params, dtype=np.float32))


def load(path: str = WEIGHTS) -> np.ndarray:
    p = np.load(path)
    if p.dtype != np.float32 or p.size != PARAMS:
        raise ValueError(f"{path}: not a router parameter file")
    return p


def train_main(pairs: int = 1_000_000, chunk: int = 125_000, epochs: int = 6, seed: int = 20261206,
               out: str = WEIGHTS):
    """Train on `pairs` routing-workload pairs labelled by the exact aligner."""
    import sys
    import time
    from . import workloads as W

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    feats, labels = [], []
    t0 = time.perf_counter()
    for k in range(0, pairs, chunk):
        ps = W.wfa_routing(n=min(chunk, pairs - k), seed=seed + k)
        f, y, b = label_batch(ps)
        feats.append(f)
        labels.append(y)
        del b
        log(f"labelled {k + ps.n} pairs ({time.perf_counter() - t0:.1f} s)")
    feat = torch.cat(feats)
    label = torch.cat(labels)
    log("label histogram:", np.bincount(label.cpu().numpy(), minlength=BUCKETS).tolist())
    mlp = Mlp(init_params(seed))
    t1 = time.perf_counter()
    hist = mlp.train(feat, label, epochs=epochs, log=log)
    torch.cuda.synchronize()
    log(f"trained {epochs} epochs in {time.perf_counter() - t1:.1f} s, loss {hist}")
    save(mlp.numpy(), out)
    log(f"saved {out}")
    return hist


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the router on the GPU")
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    train_main(pairs=a.pairs, epochs=a.epochs, out=a.out)
