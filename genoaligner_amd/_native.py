"""ctypes binding of libgenoaligner_amd.so (the C ABI in
include/genoaligner/genoaligner_c.h).

The library is required: importing a compute entry point without it raises,
there is no CPU fallback.  It is built in-tree by ``python -m
genoaligner_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

LIB_PATH = _build.LIB

GA_OK, GA_INVALID_ARGUMENT, GA_DEVICE_ERROR = 0, 1, 2
GA_MLP_PARAMS = 26896
GA_MAX_SMAX = 511
GA_MAX_SMAX_EXT = 4000
GA_SW_MAX_TRACE_CELLS = 1 << 26

# Every symbol include/genoaligner/genoaligner_c.h declares.
EXPORTS = (
    "ga_align_batch", "ga_align_batch_ext", "ga_align_sw_batch",
    "ga_backend_name", "ga_device_name", "ga_device_available",
    "ga_cigar_arena_bytes_wfa", "ga_cigar_arena_bytes_sw", "ga_sw_validate_cigar",
    "ga_wfa_device", "ga_wfa_device_noalloc", "ga_sw_device", "ga_sw_validate_device", "ga_set_kernel_timing", "ga_last_kernel_ms",
    "ga_sketch_device", "ga_mlp_forward_device", "ga_mlp_logits_device", "ga_mlp_backward_device", "ga_mlp_sgd_device",
    "ga_mlp_route_device", "ga_route_bucket_bound", "ga_wfa_routed_device", "ga_router_set", "ga_router_enabled",
    "ga_align_sw_matrix_batch", "ga_sw_matrix_blosum62", "ga_set_thread_workspace_budget",
    "ga_trim_workspace",
    "ga_wfa_graph_create", "ga_wfa_graph_launch", "ga_wfa_graph_destroy",
    "ga_wfa_autotune", "ga_wfa_set_tuning", "ga_wfa_get_tuning",
    "ga_fasta_parse", "ga_fasta_read_file", "ga_fasta_count", "ga_fasta_record", "ga_fasta_free",
    "ga_fasta_sizes", "ga_fasta_export", "ga_fasta_device", "ga_gene_qc_walk", "ga_gene_qc_tsv",
    "ga_msa_align_profiles", "ga_msa_profile_from_rows", "ga_msa_pp_device", "ga_kmer_distances", "ga_nj_tree",
)


class QcRow(C.Structure):
    _fields_ = [("cov_text", C.c_double), ("fs_runs", C.c_int32), ("fs_net", C.c_int32),
                ("stops", C.c_int32), ("stop_end", C.c_int32)]


class FastaDevOut(C.Structure):
    _fields_ = [("seq", C.c_void_p), ("seq_off", C.c_void_p), ("seq_len", C.c_void_p),
                ("acgt", C.c_void_p), ("pack2", C.c_void_p), ("max_records", C.c_int64)]


class AlignReq(C.Structure):
    _fields_ = [("pattern", C.c_void_p), ("pattern_len", C.c_int32),
                ("text", C.c_void_p), ("text_len", C.c_int32),
                ("smax", C.c_int32), ("with_cigar", C.c_int32)]


class AlignRes(C.Structure):
    _fields_ = [("score", C.c_int32), ("resolved", C.c_int32), ("cigar_off", C.c_int64),
                ("cigar_len", C.c_int32), ("rescore_ok", C.c_int32), ("wellformed_ok", C.c_int32)]


class SWScoringC(C.Structure):
    _fields_ = [("match", C.c_int32), ("mismatch", C.c_int32),
                ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class SWReq(C.Structure):
    _fields_ = [("text", C.c_void_p), ("text_len", C.c_int32),
                ("pattern", C.c_void_p), ("pattern_len", C.c_int32),
                ("scoring", SWScoringC), ("with_cigar", C.c_int32)]


class SWRes(C.Structure):
    _fields_ = [("score", C.c_int32), ("resolved", C.c_int32), ("cigar_off", C.c_int64),
                ("cigar_len", C.c_int32), ("start_i", C.c_int32), ("start_j", C.c_int32),
                ("end_i", C.c_int32), ("end_j", C.c_int32), ("rescore_ok", C.c_int32),
                ("wellformed_ok", C.c_int32), ("too_large", C.c_int32)]


class WfaDevBatch(C.Structure):
    _fields_ = [("pattern", C.c_void_p), ("pattern_off", C.c_void_p), ("pattern_len", C.c_void_p),
                ("text", C.c_void_p), ("text_off", C.c_void_p), ("text_len", C.c_void_p),
                ("smax", C.c_void_p), ("n_pairs", C.c_int64), ("smax_max", C.c_int32),
                ("max_len", C.c_int32), ("score", C.c_void_p), ("cigar", C.c_void_p),
                ("cigar_off", C.c_void_p), ("cigar_len", C.c_void_p),
                ("want_cigar", C.c_void_p), ("host_smax", C.c_void_p),
                ("host_pattern_len", C.c_void_p), ("host_text_len", C.c_void_p)]


class SwMatrix(C.Structure):
    _fields_ = [("alphabet", C.c_char * 32), ("score", (C.c_int8 * 32) * 32), ("unknown", C.c_int32),
                ("gap_open", C.c_int32), ("gap_extend", C.c_int32)]


class WfaTuning(C.Structure):
    _fields_ = [("waves", C.c_int32 * 4), ("class_ms", C.c_float * 4)]


class RouteStats(C.Structure):
    _fields_ = [("shortcut", C.c_int64), ("routed", C.c_int64), ("plain", C.c_int64), ("rerun", C.c_int64)]


class SwDevBatch(C.Structure):
    _fields_ = [("text", C.c_void_p), ("text_off", C.c_void_p), ("text_len", C.c_void_p),
                ("pattern", C.c_void_p), ("pattern_off", C.c_void_p), ("pattern_len", C.c_void_p),
                ("n_pairs", C.c_int64), ("max_text_len", C.c_int32), ("max_pattern_len", C.c_int32),
                ("scoring", SWScoringC), ("with_cigar", C.c_int32), ("out", C.c_void_p),
                ("cigar", C.c_void_p), ("cigar_off", C.c_void_p), ("cigar_len", C.c_void_p)]


class MsaParams(C.Structure):
    _fields_ = [("match", C.c_float), ("ts", C.c_float), ("tv", C.c_float),
                ("gap_open", C.c_float), ("gap_extend", C.c_float), ("free_end_gaps", C.c_int32),
                ("psgp", C.c_int32), ("psgp_scale", C.c_float), ("psgp_min_open", C.c_float),
                ("psgp_min_ext", C.c_float), ("alpha", C.c_int32), ("sub", C.c_float * 400)]


class MsaProfile(C.Structure):
    _fields_ = [("cols", C.c_void_p), ("occ", C.c_void_p), ("len", C.c_int32)]


class MsaPPRes(C.Structure):
    _fields_ = [("score", C.c_float), ("ai", C.c_int32), ("aj", C.c_int32), ("bi", C.c_int32),
                ("bj", C.c_int32), ("cigar_off", C.c_int64), ("cigar_len", C.c_int32)]


class MsaDevBatch(C.Structure):
    _fields_ = [("cols", C.c_void_p), ("occ", C.c_void_p), ("n_cols", C.c_int64), ("col_off", C.c_void_p),
                ("len", C.c_void_p), ("n_profiles", C.c_int32), ("a_idx", C.c_void_p), ("b_idx", C.c_void_p),
                ("n_pairs", C.c_int64), ("score", C.c_void_p), ("span", C.c_void_p), ("cigar", C.c_void_p),
                ("cigar_off", C.c_void_p), ("cigar_len", C.c_void_p), ("host_len", C.c_void_p),
                ("host_a_idx", C.c_void_p), ("host_b_idx", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m genoaligner_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        i32p, cpp = C.POINTER(C.c_int32), C.POINTER(C.c_char_p)
        for name, req, res in (("ga_align_batch", AlignReq, AlignRes),
                               ("ga_align_batch_ext", AlignReq, AlignRes),
                               ("ga_align_sw_batch", SWReq, SWRes)):
            f = getattr(L, name)
            f.argtypes = [C.POINTER(req), C.c_int64, C.POINTER(res), C.c_void_p, C.c_int64,
                          i32p, i32p, cpp]
            f.restype = C.c_int
        L.ga_backend_name.restype = C.c_char_p
        L.ga_device_name.restype = C.c_char_p
        L.ga_device_available.restype = C.c_int
        L.ga_cigar_arena_bytes_wfa.argtypes = [C.POINTER(AlignReq), C.c_int64]
        L.ga_cigar_arena_bytes_wfa.restype = C.c_int64
        L.ga_cigar_arena_bytes_sw.argtypes = [C.POINTER(SWReq), C.c_int64]
        L.ga_cigar_arena_bytes_sw.restype = C.c_int64
        L.ga_sw_validate_cigar.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(SWScoringC),
                                           C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.ga_sw_validate_cigar.restype = C.c_int
        L.ga_wfa_device.argtypes = [C.POINTER(WfaDevBatch), C.c_void_p, cpp]
        L.ga_wfa_device.restype = C.c_int
        L.ga_wfa_device_noalloc.argtypes = [C.POINTER(WfaDevBatch), C.c_void_p, cpp]
        L.ga_wfa_device_noalloc.restype = C.c_int
        L.ga_sw_device.argtypes = [C.POINTER(SwDevBatch), C.c_void_p, cpp]
        L.ga_sw_device.restype = C.c_int
        L.ga_sw_validate_device.argtypes = [C.POINTER(SwDevBatch), C.c_void_p, C.c_void_p, cpp]
        L.ga_sw_validate_device.restype = C.c_int
        L.ga_set_kernel_timing.argtypes = [C.c_int]
        L.ga_set_kernel_timing.restype = None
        L.ga_last_kernel_ms.argtypes = []
        L.ga_last_kernel_ms.restype = C.c_float
        vp, i64, f32 = C.c_void_p, C.c_int64, C.c_float
        L.ga_sketch_device.argtypes = [C.POINTER(WfaDevBatch), vp, vp, vp, cpp]
        L.ga_mlp_forward_device.argtypes = [vp, vp, i64, vp, vp, vp, vp, cpp]
        L.ga_mlp_logits_device.argtypes = [vp, vp, i64, vp, vp, cpp]
        L.ga_mlp_backward_device.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, cpp]
        L.ga_mlp_sgd_device.argtypes = [vp, vp, vp, f32, f32, vp, cpp]
        L.ga_mlp_route_device.argtypes = [vp, i64, f32, vp, vp, cpp]
        L.ga_wfa_routed_device.argtypes = [C.POINTER(WfaDevBatch), vp, f32, C.POINTER(RouteStats), vp, cpp]
        for f in (L.ga_sketch_device, L.ga_mlp_forward_device, L.ga_mlp_logits_device, L.ga_mlp_backward_device,
                  L.ga_mlp_sgd_device, L.ga_mlp_route_device, L.ga_wfa_routed_device):
            f.restype = C.c_int
        L.ga_router_set.argtypes = [vp, f32, cpp]
        L.ga_router_set.restype = C.c_int
        L.ga_router_enabled.argtypes = []
        L.ga_router_enabled.restype = C.c_int
        L.ga_align_sw_matrix_batch.argtypes = [C.POINTER(SWReq), C.c_int64, C.POINTER(SwMatrix),
                                               C.POINTER(SWRes), C.c_void_p, C.c_int64, i32p, i32p, cpp]
        L.ga_align_sw_matrix_batch.restype = C.c_int
        L.ga_sw_matrix_blosum62.argtypes = [C.POINTER(SwMatrix), C.c_int32, C.c_int32]
        L.ga_sw_matrix_blosum62.restype = None
        L.ga_wfa_autotune.argtypes = [C.POINTER(WfaDevBatch), C.c_int, C.POINTER(WfaTuning), C.c_void_p, cpp]
        L.ga_wfa_autotune.restype = C.c_int
        L.ga_wfa_set_tuning.argtypes = [C.POINTER(WfaTuning)]
        L.ga_wfa_set_tuning.restype = None
        L.ga_wfa_get_tuning.argtypes = [C.POINTER(WfaTuning)]
        L.ga_wfa_get_tuning.restype = None
        L.ga_wfa_graph_create.argtypes = [C.POINTER(WfaDevBatch), C.POINTER(C.c_void_p), cpp]
        L.ga_wfa_graph_create.restype = C.c_int
        L.ga_wfa_graph_launch.argtypes = [C.c_void_p, C.c_void_p, cpp]
        L.ga_wfa_graph_launch.restype = C.c_int
        L.ga_wfa_graph_destroy.argtypes = [C.c_void_p]
        L.ga_wfa_graph_destroy.restype = None
        L.ga_set_thread_workspace_budget.argtypes = [C.c_int64]
        L.ga_set_thread_workspace_budget.restype = None
        L.ga_trim_workspace.argtypes = []
        L.ga_trim_workspace.restype = None
        L.ga_fasta_parse.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_void_p), cpp]
        L.ga_fasta_parse.restype = C.c_int
        L.ga_fasta_read_file.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), cpp]
        L.ga_fasta_read_file.restype = C.c_int
        L.ga_fasta_count.argtypes = [C.c_void_p]
        L.ga_fasta_count.restype = C.c_int64
        i64p = C.POINTER(C.c_int64)
        L.ga_fasta_record.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p), i64p,
                                      C.POINTER(C.c_void_p), i64p, C.POINTER(C.c_void_p), i64p]
        L.ga_fasta_record.restype = C.c_int
        L.ga_fasta_sizes.argtypes = [C.c_void_p, i64p, i64p, i64p]
        L.ga_fasta_sizes.restype = None
        L.ga_fasta_export.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.ga_fasta_export.restype = C.c_int
        L.ga_fasta_free.argtypes = [C.c_void_p]
        L.ga_fasta_free.restype = None
        L.ga_fasta_device.argtypes = [C.c_void_p, C.c_int64, C.POINTER(FastaDevOut), i64p, i64p,
                                      C.c_void_p, cpp]
        L.ga_fasta_device.restype = C.c_int
        L.ga_gene_qc_walk.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int32, C.c_int32, C.c_void_p]
        L.ga_gene_qc_walk.restype = C.c_int
        L.ga_gene_qc_tsv.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p]
        L.ga_gene_qc_tsv.restype = C.c_int64
        L.ga_msa_align_profiles.argtypes = [C.POINTER(MsaProfile), C.POINTER(MsaProfile), C.c_int64,
                                            C.POINTER(MsaParams), C.POINTER(MsaPPRes), C.c_void_p, C.c_int64, cpp]
        L.ga_msa_align_profiles.restype = C.c_int
        L.ga_msa_profile_from_rows.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
        L.ga_msa_profile_from_rows.restype = C.c_int
        L.ga_msa_pp_device.argtypes = [C.POINTER(MsaDevBatch), C.POINTER(MsaParams), C.c_void_p, cpp]
        L.ga_msa_pp_device.restype = C.c_int
        L.ga_kmer_distances.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_void_p, cpp]
        L.ga_kmer_distances.restype = C.c_int
        L.ga_nj_tree.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, cpp]
        L.ga_nj_tree.restype = C.c_int
        L.ga_route_bucket_bound.argtypes = [C.c_int32]
        L.ga_route_bucket_bound.restype = C.c_int32
        _lib = L
    return _lib
