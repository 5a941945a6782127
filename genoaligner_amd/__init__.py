"""genoaligner_amd -- genoaligner's batched pairwise alignment path (WFA edit
distance and affine Smith-Waterman) on AMD Instinct MI355X (gfx950).

Public surface mirrors the reference's ``genoaligner::`` API; see
``genoaligner_amd.api``.  Device-resident batches (sequences already in HBM)
go through ``genoaligner_amd.device``.
"""
from .api import (  # noqa: F401
    AlignRequest, AlignResult, BatchResult, SWAlignResult, SWBatchResult, SWRequest,
    SWScoring, Status, SW_MAX_TRACE_CELLS, align, align_batch, align_batch_ext, align_sw,
    align_sw_batch, align_sw_matrix_batch, backend_name, blosum62, device_available, device_name,
    router_enabled, set_router,
)
