"""Device-resident batches: sequences and results live in HBM (torch tensors
as plain device memory), work is enqueued on a HIP stream through the C ABI
entry points ga_wfa_device / ga_sw_device.

Used by the benchmark (inputs resident before the timed region) and by the GPU
parity tests.  Nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _native as N
from .workloads import PairSet

PAD = 128   # readable bytes before the first / after the last sequence


def _stream_handle(stream: Optional[torch.cuda.Stream]) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _packed(buf: np.ndarray, off: np.ndarray, dev):
    padded = np.zeros(len(buf) + 2 * PAD, dtype=np.uint8)
    padded[PAD: PAD + len(buf)] = buf
    return (torch.from_numpy(padded).to(dev),
            torch.from_numpy((off + PAD).astype(np.int64)).to(dev))


def _check(rc: int, err: C.c_char_p, what: str):
    if rc != N.GA_OK:
        raise RuntimeError(f"{what} failed ({rc}): {err.value.decode() if err.value else '?'}")


class WfaDeviceBatch:
    """WFA over a PairSet (a = pattern, b = text) with per-pair smax.

    want: optional per-pair CIGAR flags (0 = score only for that pair; needs
    trace).  classes: hand the host copies of smax and the lengths to the
    library so it splits the batch into smax size classes, longest pair first.
    """

    def __init__(self, ps: PairSet, trace: bool = True, device: str = "cuda:0",
                 want: Optional[np.ndarray] = None, classes: bool = True):
        dev = torch.device(device)
        self.n = ps.n
        self.trace = trace
        self.pat, self.poff = _packed(ps.a, ps.a_off, dev)
        self.txt, self.toff = _packed(ps.b, ps.b_off, dev)
        self.plen = torch.from_numpy(ps.a_len.astype(np.int32)).to(dev)
        self.tlen = torch.from_numpy(ps.b_len.astype(np.int32)).to(dev)
        self.smax = torch.from_numpy(ps.smax.astype(np.int32)).to(dev)
        self.score = torch.empty(self.n, dtype=torch.int32, device=dev)
        slot = ps.a_len.astype(np.int64) + ps.b_len.astype(np.int64)
        self.slot = slot
        self.slot_off_np = np.zeros(self.n, dtype=np.int64)
        if self.n > 1:
            self.slot_off_np[1:] = np.cumsum(slot)[:-1]
        self.cig_bytes = int(slot.sum())
        if trace:
            self.cigar = torch.empty(max(self.cig_bytes, 1), dtype=torch.uint8, device=dev)
            self.cig_off = torch.from_numpy(self.slot_off_np).to(dev)
            self.cig_len = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.smax_max = int(ps.smax.max()) if self.n else 0
        self.max_len = int(max(ps.a_len.max(), ps.b_len.max())) if self.n else 0
        self.want_np = None
        if want is not None:
            assert trace and len(want) == self.n
            self.want_np = np.ascontiguousarray(want, dtype=np.uint8)
            self.want = torch.from_numpy(self.want_np).to(dev)
        # host copies stay referenced by self for as long as the batch lives
        self.h_smax = np.ascontiguousarray(ps.smax, dtype=np.int32)
        self.h_plen = np.ascontiguousarray(ps.a_len, dtype=np.int32)
        self.h_tlen = np.ascontiguousarray(ps.b_len, dtype=np.int32)
        hp = (lambda a: a.ctypes.data) if classes else (lambda a: None)
        self.b = N.WfaDevBatch(
            self.pat.data_ptr(), self.poff.data_ptr(), self.plen.data_ptr(),
            self.txt.data_ptr(), self.toff.data_ptr(), self.tlen.data_ptr(),
            self.smax.data_ptr(), self.n, self.smax_max, self.max_len, self.score.data_ptr(),
            self.cigar.data_ptr() if trace else None, self.cig_off.data_ptr() if trace else None,
            self.cig_len.data_ptr() if trace else None,
            self.want.data_ptr() if want is not None else None,
            hp(self.h_smax), hp(self.h_plen), hp(self.h_tlen))

    def run(self, stream: Optional[torch.cuda.Stream] = None):
        err = C.c_char_p()
        _check(N.lib().ga_wfa_device(C.byref(self.b), C.c_void_p(_stream_handle(stream)),
                                     C.byref(err)), err, "ga_wfa_device")

    def graph(self) -> "WfaGraph":
        """Capture the device-planned submission of this batch into a HIP graph
        (its buffers are baked in: refill them in place and replay)."""
        return WfaGraph(self)

    def results(self):
        """(scores, cigars) on the host; cigars[i] is bytes ('' when none)."""
        torch.cuda.synchronize()
        sc = self.score.cpu().numpy()
        if not self.trace:
            return sc, None
        ln = self.cig_len.cpu().numpy()
        raw = self.cigar.cpu().numpy().tobytes()
        cig = []
        for i in range(self.n):
            end = int(self.slot_off_np[i] + self.slot[i])
            L = int(ln[i]) if sc[i] >= 0 else 0
            if self.want_np is not None and not self.want_np[i]:
                L = 0
            cig.append(raw[end - L: end])
        return sc, cig


def wfa_autotune(batch: "WfaDeviceBatch", reps: int = 2, apply: bool = True):
    """Time waves-per-workgroup candidates per WFA size class on `batch` (needs
    classes=True) and, with apply, make the fastest current.  Returns
    (waves per class, ms per class)."""
    t = N.WfaTuning()
    err = C.c_char_p()
    _check(N.lib().ga_wfa_autotune(C.byref(batch.b), int(reps), C.byref(t), C.c_void_p(_stream_handle(None)),
                                   C.byref(err)), err, "ga_wfa_autotune")
    if apply:
        N.lib().ga_wfa_set_tuning(C.byref(t))
    return list(t.waves), list(t.class_ms)


def wfa_set_tuning(waves=None):
    """Waves per workgroup for the four size classes (None: defaults)."""
    if waves is None:
        N.lib().ga_wfa_set_tuning(None)
        return
    t = N.WfaTuning()
    for c in range(4):
        t.waves[c] = int(waves[c])
    N.lib().ga_wfa_set_tuning(C.byref(t))


def wfa_get_tuning():
    t = N.WfaTuning()
    N.lib().ga_wfa_get_tuning(C.byref(t))
    return list(t.waves)


class WfaGraph:
    def __init__(self, batch: WfaDeviceBatch):
        self.batch = batch                      # keeps the buffers alive
        h = C.c_void_p()
        err = C.c_char_p()
        _check(N.lib().ga_wfa_graph_create(C.byref(batch.b), C.byref(h), C.byref(err)), err,
               "ga_wfa_graph_create")
        self.h = h

    def launch(self, stream: Optional[torch.cuda.Stream] = None):
        err = C.c_char_p()
        _check(N.lib().ga_wfa_graph_launch(self.h, C.c_void_p(_stream_handle(stream)), C.byref(err)), err,
               "ga_wfa_graph_launch")

    def close(self):
        if self.h:
            N.lib().ga_wfa_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SwDeviceBatch:
    """Smith-Waterman over a PairSet (a = text rows, b = pattern columns)."""

    def __init__(self, ps: PairSet, scoring=(2, -3, 3, 1), trace: bool = True,
                 device: str = "cuda:0"):
        dev = torch.device(device)
        self.n = ps.n
        self.trace = trace
        self.txt, self.toff = _packed(ps.a, ps.a_off, dev)
        self.pat, self.poff = _packed(ps.b, ps.b_off, dev)
        self.tlen = torch.from_numpy(ps.a_len.astype(np.int32)).to(dev)
        self.plen = torch.from_numpy(ps.b_len.astype(np.int32)).to(dev)
        self.out = torch.empty(self.n * 5, dtype=torch.int32, device=dev)
        slot = ps.a_len.astype(np.int64) + ps.b_len.astype(np.int64)
        self.slot = slot
        self.slot_off_np = np.zeros(self.n, dtype=np.int64)
        if self.n > 1:
            self.slot_off_np[1:] = np.cumsum(slot)[:-1]
        if trace:
            self.cigar = torch.empty(max(int(slot.sum()), 1), dtype=torch.uint8, device=dev)
            self.cig_off = torch.from_numpy(self.slot_off_np).to(dev)
            self.cig_len = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.b = N.SwDevBatch(
            self.txt.data_ptr(), self.toff.data_ptr(), self.tlen.data_ptr(),
            self.pat.data_ptr(), self.poff.data_ptr(), self.plen.data_ptr(), self.n,
            int(ps.a_len.max()) if self.n else 0, int(ps.b_len.max()) if self.n else 0,
            N.SWScoringC(*scoring), 1 if trace else 0, self.out.data_ptr(),
            self.cigar.data_ptr() if trace else None, self.cig_off.data_ptr() if trace else None,
            self.cig_len.data_ptr() if trace else None)

    def run(self, stream: Optional[torch.cuda.Stream] = None):
        err = C.c_char_p()
        _check(N.lib().ga_sw_device(C.byref(self.b), C.c_void_p(_stream_handle(stream)),
                                    C.byref(err)), err, "ga_sw_device")

    def validate(self, stream: Optional[torch.cuda.Stream] = None) -> np.ndarray:
        """The CIGAR checks on the device (ga_sw_validate_device) over the
        batch's current results: per pair rescore_ok | wellformed_ok << 1 (0 for
        pairs with score 0)."""
        if not self.trace:
            raise ValueError("a score-only batch has no CIGARs to check")
        flags = torch.empty(max(self.n, 1), dtype=torch.uint8, device=self.out.device)
        err = C.c_char_p()
        _check(N.lib().ga_sw_validate_device(C.byref(self.b), C.c_void_p(flags.data_ptr()),
                                             C.c_void_p(_stream_handle(stream)), C.byref(err)),
               err, "ga_sw_validate_device")
        torch.cuda.synchronize()
        return flags[:self.n].cpu().numpy()

    def results(self):
        """(out[n,5] = score, start_i, start_j, end_i, end_j; cigars or None)."""
        torch.cuda.synchronize()
        out = self.out.cpu().numpy().reshape(self.n, 5)
        if not self.trace:
            return out, None
        ln = self.cig_len.cpu().numpy()
        raw = self.cigar.cpu().numpy().tobytes()
        cig = []
        for i in range(self.n):
            end = int(self.slot_off_np[i] + self.slot[i])
            L = int(ln[i]) if out[i, 0] > 0 else 0
            cig.append(raw[end - L: end])
        return out, cig
