"""MI355X-native ordinal retrieval (ordvec's RankQuant / SignBitmap /
FastScan hot path, with the Rank and Bitmap indexes) behind ordvec's C ABI.

Importing the package loads ``ordvec_amd/lib/libordvec_amd.so``; there is no
other compute path.
"""
from ._native import OrdvecError, load_library  # noqa: F401

load_library()

from .index import (  # noqa: E402,F401
    MAX_DIM,
    MAX_SIGN_BITMAP_DIM,
    MAX_VECTORS,
    Bitmap,
    FastscanBatch,
    Rank,
    RankQuant,
    RankQuantFastscan,
    SignBitmap,
    TwoStageBatch,
    TwoStageCandidatePolicy,
    stage1_exact_queries,
    synchronize,
    two_stage_add,
    two_stage_search,
    two_stage_search_policy,
)
from .primitives import (  # noqa: E402,F401
    bucket_centre,
    bucket_ranks,
    pack_buckets,
    rank_norm,
    rank_to_bucket,
    rank_transform,
    rankquant_bytes_per_vec,
    rankquant_eval_search,
    rankquant_norm,
    unpack_buckets,
)
from .serving import Server  # noqa: E402,F401

try:  # the autograd binding needs torch
    from .autograd import rerank_scores  # noqa: E402,F401
except ImportError:  # pragma: no cover
    pass

__all__ = [
    "Rank",
    "RankQuant",
    "Bitmap",
    "FastscanBatch",
    "SignBitmap",
    "RankQuantFastscan",
    "two_stage_add",
    "two_stage_search",
    "two_stage_search_policy",
    "TwoStageCandidatePolicy",
    "TwoStageBatch",
    "Server",
    "rerank_scores",
    "stage1_exact_queries",
    "synchronize",
    "OrdvecError",
    "MAX_DIM",
    "MAX_SIGN_BITMAP_DIM",
    "MAX_VECTORS",
    "rank_transform",
    "rank_to_bucket",
    "bucket_ranks",
    "pack_buckets",
    "unpack_buckets",
    "rankquant_bytes_per_vec",
    "bucket_centre",
    "rank_norm",
    "rankquant_norm",
    "rankquant_eval_search",
]
