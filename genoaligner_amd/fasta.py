"""FASTA reading (reference: include/genoaligner/io/fasta.hpp:38-175), over the
C ABI of libgenoaligner_amd.so.

``read_fasta`` / ``read_fasta_file`` mirror ``io::read_fasta`` /
``io::read_fasta_file``: same records (first header token the id, the rest the
description, sequence bytes kept literally with the line breaks removed) and
the same error messages.  Where the reference returns ``false`` and fills
``*error``, these raise ``FastaError`` carrying that message; nothing partial
is returned.

``fasta_to_device`` is the device ingest: a whole FASTA file in HBM becomes the
back-to-back sequence bytes the batch kernels read, per-record offsets and
lengths, an A/C/G/T flag per record and the 2-bit codes of the stream.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Union

from . import _native as N


class FastaError(ValueError):
    """A malformed or unreadable FASTA file; str() is the reference's message."""


@dataclass
class FastaRecord:
    id: bytes
    description: bytes
    sequence: bytes


@dataclass
class FastaArrays:
    """Every record of a file as three byte buffers with offsets (n + 1 each):
    record k's id is ids[id_off[k]:id_off[k + 1]], likewise desc and seq; each
    buffer holds one zero byte behind its last record."""
    ids: "np.ndarray"
    id_off: "np.ndarray"
    desc: "np.ndarray"
    desc_off: "np.ndarray"
    seq: "np.ndarray"
    seq_off: "np.ndarray"

    def __len__(self) -> int:
        return len(self.id_off) - 1


def _export(h: C.c_void_p) -> FastaArrays:
    """Every record, through one bulk copy (ga_fasta_sizes / ga_fasta_export)."""
    import numpy as np
    L = N.lib()
    n = L.ga_fasta_count(h)
    sz = [C.c_int64() for _ in range(3)]
    L.ga_fasta_sizes(h, C.byref(sz[0]), C.byref(sz[1]), C.byref(sz[2]))
    bufs = [np.empty(x.value + 1, np.uint8) for x in sz]
    for b in bufs:
        b[-1] = 0                                # one zero byte behind the last record
    offs = [np.empty(n + 1, np.int64) for _ in range(3)]
    L.ga_fasta_export(h, bufs[0].ctypes.data, offs[0].ctypes.data, bufs[1].ctypes.data, offs[1].ctypes.data,
                      bufs[2].ctypes.data, offs[2].ctypes.data)
    return FastaArrays(bufs[0], offs[0], bufs[1], offs[1], bufs[2], offs[2])


def _records(a: FastaArrays) -> List[FastaRecord]:
    raw = [b.tobytes() for b in (a.ids, a.desc, a.seq)]
    o = [x.tolist() for x in (a.id_off, a.desc_off, a.seq_off)]
    return [FastaRecord(raw[0][o[0][k]:o[0][k + 1]], raw[1][o[1][k]:o[1][k + 1]], raw[2][o[2][k]:o[2][k + 1]])
            for k in range(len(a))]


def _take_arrays(rc: int, h: C.c_void_p, err: C.c_char_p) -> FastaArrays:
    if rc != N.GA_OK:
        raise FastaError((err.value or b"").decode("latin-1"))
    try:
        return _export(h)
    finally:
        N.lib().ga_fasta_free(h)


def _take(rc: int, h: C.c_void_p, err: C.c_char_p) -> List[FastaRecord]:
    return _records(_take_arrays(rc, h, err))


def read_fasta_file_arrays(path: Union[str, os.PathLike]) -> FastaArrays:
    """read_fasta_file's records as byte buffers and offsets, with no Python
    object per record (for callers that hand the sequences to a batch)."""
    h, err = C.c_void_p(), C.c_char_p()
    return _take_arrays(N.lib().ga_fasta_read_file(os.fsencode(path), C.byref(h), C.byref(err)), h, err)


def read_fasta(data: Union[bytes, bytearray, memoryview, str]) -> List[FastaRecord]:
    """Every record of an in-memory FASTA file (fasta.hpp:59-157)."""
    if isinstance(data, str):
        data = data.encode("latin-1")
    data = bytes(data)
    h, err = C.c_void_p(), C.c_char_p()
    return _take(N.lib().ga_fasta_parse(data, len(data), C.byref(h), C.byref(err)), h, err)


def read_fasta_file(path: Union[str, os.PathLike]) -> List[FastaRecord]:
    """Every record of a FASTA file on disk (fasta.hpp:159-170)."""
    h, err = C.c_void_p(), C.c_char_p()
    return _take(N.lib().ga_fasta_read_file(os.fsencode(path), C.byref(h), C.byref(err)), h, err)


@dataclass
class DeviceFasta:
    """Sequences of a FASTA file in HBM: record k is seq[seq_off[k]:][:seq_len[k]];
    acgt[k] == 1 iff all its bytes are A, C, G or T; pack2 (int32 words, read as
    uint32) holds 2-bit codes of all of seq, byte p at bits 2*(p % 16) of word
    p // 16, A 0, C 1, G 2, T 3."""
    n: int
    total: int
    seq: "torch.Tensor"
    seq_off: "torch.Tensor"
    seq_len: "torch.Tensor"
    acgt: "torch.Tensor"
    pack2: Optional["torch.Tensor"]


def fasta_to_device(data, stream=None, pack2: bool = True) -> DeviceFasta:
    """Parse a FASTA file on the device.  ``data``: a uint8 CUDA tensor holding
    the file (or bytes, copied to the current device first).  Raises
    ``FastaError`` with the host reader's message for a malformed file."""
    import torch
    if not isinstance(data, torch.Tensor):
        raw = bytes(data.encode("latin-1") if isinstance(data, str) else data)
        data = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda() if raw else \
            torch.empty(0, dtype=torch.uint8, device="cuda")
    if data.dtype != torch.uint8 or not data.is_cuda or data.dim() != 1 or not data.is_contiguous():
        raise ValueError("data must be a contiguous 1-D uint8 CUDA tensor")
    dev = data.device
    ln = data.numel()
    st = torch.cuda.current_stream(dev) if stream is None else stream
    with torch.cuda.stream(st):
        seq = torch.empty(max(ln, 1), dtype=torch.uint8, device=dev)
        pk = torch.empty(max((ln + 15) // 16, 1), dtype=torch.int32, device=dev) if pack2 else None
    # per-record arrays sized for records of 64 bytes on average; the call
    # reports the count it needs when the file holds more, and runs again
    cap = max(16, ln // 64)
    while True:
        with torch.cuda.stream(st):
            off = torch.empty(cap, dtype=torch.int64, device=dev)
            lens = torch.empty(cap, dtype=torch.int32, device=dev)
            acgt = torch.empty(cap, dtype=torch.uint8, device=dev)
        o = N.FastaDevOut(seq.data_ptr(), off.data_ptr(), lens.data_ptr(), acgt.data_ptr(),
                          pk.data_ptr() if pk is not None else None, cap)
        n, tot, err = C.c_int64(), C.c_int64(), C.c_char_p()
        rc = N.lib().ga_fasta_device(C.c_void_p(data.data_ptr() if ln else None), ln, C.byref(o),
                                     C.byref(n), C.byref(tot), C.c_void_p(st.cuda_stream), C.byref(err))
        if rc == N.GA_INVALID_ARGUMENT and n.value > cap:
            cap = n.value
            continue
        break
    if rc == N.GA_INVALID_ARGUMENT:
        raise FastaError((err.value or b"").decode("latin-1"))
    if rc != N.GA_OK:
        raise RuntimeError(f"ga_fasta_device: {(err.value or b'').decode()}")
    k, t = n.value, tot.value
    return DeviceFasta(k, t, seq[:t], off[:k], lens[:k], acgt[:k],
                       pk[:(t + 15) // 16] if pk is not None else None)
