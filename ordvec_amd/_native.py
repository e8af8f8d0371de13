"""ctypes binding to ``libordvec_amd.so`` (the ordvec C ABI, ``include/ordvec.h``).

The library is the only compute path: if it is missing this module raises at
import time instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

# ORDVEC_AMD_LIB names another build of the same library (the CPU sanitizer
# build of the host code is one).
LIB_PATH = os.environ.get("ORDVEC_AMD_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                                                            "libordvec_amd.so")

# Status codes (include/ordvec.h).
OK = 0
NULL_POINTER = 1
BAD_ARGUMENT = 2
BAD_STRUCT_SIZE = 3
UNSUPPORTED_FORMAT = 4
CORRUPT_INDEX = 5
IO = 6
DIM_MISMATCH = 7
NONFINITE_QUERY = 8
ROW_ID_OUT_OF_RANGE = 9
BUFFER_TOO_SMALL = 10
UNSUPPORTED_OPERATION = 11
PANIC = 12
INTERNAL = 13

FLAG_DEVICE_PTRS = 1
FLAG_ASSUME_FINITE = 2
FLAG_NO_GRAPH = 4

MAX_TOPK = 1024
MAX_STAGE1_M = 4096
MAX_ENCODE_DIM = 4096
MAX_SIGN_DIM = 8192


class OrdvecError(RuntimeError):
    """A non-OK status crossing the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class OrdvecValueError(OrdvecError, ValueError):
    pass


class OrdvecIOError(OrdvecError, OSError):
    pass


class OrdvecUnsupported(OrdvecError, NotImplementedError):
    pass


class IndexInfo(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint64), ("kind", C.c_uint32), ("format_version", C.c_uint32),
        ("dim", C.c_uint64), ("bit_width", C.c_uint32), ("n_top", C.c_uint32),
        ("vector_count", C.c_uint64), ("bytes_per_vec", C.c_uint64),
        ("source_file_size_bytes", C.c_uint64), ("capabilities", C.c_uint64),
        ("reserved", C.c_uint64 * 8),
    ]


class SearchParams(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint64), ("query", C.POINTER(C.c_float)), ("dim", C.c_uint64),
        ("k", C.c_uint64), ("candidate_rows", C.POINTER(C.c_uint32)), ("candidate_count", C.c_uint64),
        ("flags", C.c_uint64), ("user_tag", C.c_uint64), ("reserved", C.c_uint64 * 8),
    ]


class SearchStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint64), ("abi_version", C.c_uint32), ("kind", C.c_uint32),
        ("dim", C.c_uint64), ("bit_width", C.c_uint32), ("n_top", C.c_uint32), ("k", C.c_uint64),
        ("user_tag", C.c_uint64), ("vector_count", C.c_uint64), ("candidate_count", C.c_uint64),
        ("returned_count", C.c_uint64), ("total_ns", C.c_uint64), ("prepare_ns", C.c_uint64),
        ("score_ns", C.c_uint64), ("select_ns", C.c_uint64), ("vectors_scored", C.c_uint64),
        ("bytes_read", C.c_uint64), ("reserved", C.c_uint64 * 8),
    ]


class ServeReport(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint64),
        ("requests", C.c_uint64),
        ("queries", C.c_uint64),
        ("errors", C.c_uint64),
        ("duration_s", C.c_double),
        ("offered_qps", C.c_double),
        ("achieved_qps", C.c_double),
        ("achieved_rps", C.c_double),
        ("p50_ms", C.c_double),
        ("p90_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("p999_ms", C.c_double),
        ("max_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("batches", C.c_double),
        ("mean_batch_queries", C.c_double),
        ("p99_queue_ms", C.c_double),
        ("p99_service_ms", C.c_double),
        ("exact_path_queries", C.c_uint64),
        ("pool_bytes", C.c_uint64),
        ("peak_pool_bytes", C.c_uint64),
        ("overlapped_batches", C.c_uint64),
        ("max_batches_in_flight", C.c_uint64),
        ("workspace_overflows", C.c_uint64),
    ]


class TwoStagePolicy(C.Structure):
    _fields_ = [("min_candidates", C.c_uint64), ("k_multiplier", C.c_uint64), ("max_candidates", C.c_uint64),
                ("has_max_candidates", C.c_uint32), ("reserved", C.c_uint32)]


class TwoStageShape(C.Structure):
    _fields_ = [("scan_tiles", C.c_uint32), ("scan_pairs", C.c_uint32), ("scan_flat", C.c_uint32),
                ("sample_wgs", C.c_uint32), ("side_prep", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class Hit(C.Structure):
    _fields_ = [("row_id", C.c_uint64), ("id", C.c_uint64), ("score", C.c_float), ("reserved", C.c_uint32)]


P = C.c_void_p
U64 = C.c_uint64
U32 = C.c_uint32
ST = C.c_uint32

# name -> (restype, argtypes); every symbol include/ordvec.h declares.
SIGNATURES = {
    "ordvec_abi_version": (U32, []),
    "ordvec_version_string": (C.c_char_p, []),
    "ordvec_status_name": (C.c_char_p, [U32]),
    "ordvec_last_error": (C.c_char_p, []),
    "ordvec_init": (ST, []),
    "ordvec_index_info_init": (None, [P]),
    "ordvec_search_params_init": (None, [P]),
    "ordvec_search_stats_init": (None, [P]),
    "ordvec_index_load": (ST, [C.c_char_p, U64, C.POINTER(P)]),
    "ordvec_index_probe": (ST, [C.c_char_p, U64, P]),
    "ordvec_index_info": (ST, [P, P]),
    "ordvec_index_free": (None, [P]),
    "ordvec_index_search": (ST, [P, P, P, U64, P, P]),
    "ordvec_set_stream": (ST, [P]),
    "ordvec_synchronize": (ST, []),
    "ordvec_stage1_exact_queries": (ST, [P]),
    "ordvec_kernel_timing_enable": (ST, [C.c_int]),
    "ordvec_engine_option": (ST, [U32, C.c_int64, C.POINTER(C.c_int64)]),
    "ordvec_kernel_timing_enable_families": (ST, [U32]),
    "ordvec_kernel_timing_read": (ST, [U32, P, P]),
    "ordvec_rankquant_new": (ST, [U64, U32, C.POINTER(P)]),
    "ordvec_rankquant_add": (ST, [P, P, U64, U64]),
    "ordvec_rankquant_info": (ST, [P, P, P, P]),
    "ordvec_rankquant_search_asymmetric": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_rankquant_search_subset": (ST, [P, P, P, U64, U64, P, P, P, U64]),
    "ordvec_rankquant_search_subset_batched": (ST, [P, P, U64, P, P, U64, U64, P, P, U64]),
    "ordvec_rankquant_codes": (ST, [P, P, U64, U64]),
    "ordvec_rankquant_eval_search": (ST, [P, U64, P, U64, U64, U32, U64, P, P, U64]),
    "ordvec_rank_transform": (ST, [P, U64, U64, P, U64]),
    "ordvec_rankquant_write": (ST, [P, C.c_char_p]),
    "ordvec_rankquant_load": (ST, [C.c_char_p, C.POINTER(P)]),
    "ordvec_rankquant_free": (None, [P]),
    "ordvec_sign_bitmap_new": (ST, [U64, C.POINTER(P)]),
    "ordvec_sign_bitmap_add": (ST, [P, P, U64, U64]),
    "ordvec_sign_bitmap_info": (ST, [P, P, P]),
    "ordvec_sign_bitmap_top_m": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_sign_bitmap_words": (ST, [P, P, U64, U64]),
    "ordvec_sign_bitmap_query_bits": (ST, [P, P, U64, P, U64]),
    "ordvec_sign_bitmap_score_all": (ST, [P, P, U64, P, U64]),
    "ordvec_sign_bitmap_write": (ST, [P, C.c_char_p]),
    "ordvec_sign_bitmap_load": (ST, [C.c_char_p, C.POINTER(P)]),
    "ordvec_sign_bitmap_free": (None, [P]),
    "ordvec_two_stage_search": (ST, [P, P, P, U64, U64, U64, P, P, U64]),
    "ordvec_two_stage_add": (ST, [P, P, P, U64, U64]),
    "ordvec_two_stage_search_noalloc": (ST, [P, P, P, U64, U64, U64, P, P, U64]),
    "ordvec_two_stage_policy_default": (None, [P]),
    "ordvec_two_stage_policy_candidate_count": (U64, [P, U64, U64]),
    "ordvec_rankquant_search_with_sign_probe": (ST, [P, P, P, U64, U64, P, P, P, P, U64]),
    "ordvec_two_stage_search_policy": (ST, [P, P, P, U64, U64, P, P, P, U64]),
    "ordvec_bitmap_new": (ST, [U64, U64, C.POINTER(P)]),
    "ordvec_bitmap_add": (ST, [P, P, U64, U64]),
    "ordvec_bitmap_info": (ST, [P, P, P, P]),
    "ordvec_bitmap_search": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_bitmap_top_m": (ST, [P, P, U64, U64, P, U64]),
    "ordvec_bitmap_search_subset_batched": (ST, [P, P, U64, P, P, U64, U64, P, P, P, U64]),
    "ordvec_bitmap_query_bits": (ST, [P, P, U64, P, U64]),
    "ordvec_bitmap_overlap_subset": (ST, [P, P, P, U64, P]),
    "ordvec_bitmap_words": (ST, [P, P, U64]),
    "ordvec_bitmap_write": (ST, [P, C.c_char_p]),
    "ordvec_bitmap_load": (ST, [C.c_char_p, C.POINTER(P)]),
    "ordvec_bitmap_free": (None, [P]),
    "ordvec_fastscan_new": (ST, [U64, C.POINTER(P)]),
    "ordvec_fastscan_add": (ST, [P, P, U64, U64]),
    "ordvec_fastscan_search": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_fastscan_info": (ST, [P, P, P]),
    "ordvec_fastscan_write": (ST, [P, C.c_char_p]),
    "ordvec_fastscan_load": (ST, [C.c_char_p, C.POINTER(P)]),
    "ordvec_fastscan_free": (None, [P]),
    "ordvec_fastscan_new_bits": (ST, [U64, U32, C.POINTER(P)]),
    "ordvec_fastscan_bits": (ST, [P, P]),
    "ordvec_fastscan_batch_new": (ST, [P, U64, P, P, C.POINTER(P)]),
    "ordvec_fastscan_batch_autotune": (ST, [P, P, U32, U64]),
    "ordvec_fastscan_batch_run": (ST, [P, P, P, P, U64]),
    "ordvec_fastscan_batch_config": (ST, [P, P, P, P, P, P, P, P]),
    "ordvec_fastscan_batch_tuning": (ST, [P, U64, P, P, P, P]),
    "ordvec_fastscan_batch_set_config": (ST, [P, U32, U32, U32]),
    "ordvec_fastscan_batch_free": (None, [P]),
    "ordvec_two_stage_batch_new": (ST, [P, P, U64, U64, U64, C.POINTER(P)]),
    "ordvec_two_stage_batch_run": (ST, [P, P, U64, P, P, U64]),
    "ordvec_two_stage_batch_autotune": (ST, [P, P, U64, U32, U64]),
    "ordvec_two_stage_batch_shape": (ST, [P, C.POINTER(TwoStageShape)]),
    "ordvec_two_stage_batch_set_shape": (ST, [P, C.POINTER(TwoStageShape)]),
    "ordvec_two_stage_batch_tuning": (ST, [P, U64, C.POINTER(TwoStageShape), P]),
    "ordvec_two_stage_batch_stats": (ST, [P, P, P, P, P, P]),
    "ordvec_two_stage_batch_free": (None, [P]),
    "ordvec_rankquant_score_ids": (ST, [P, P, U64, P, U64, P, U64]),
    "ordvec_rankquant_score_ids_backward": (ST, [P, P, U64, P, U64, P, P, U64]),
    "ordvec_server_start": (ST, [P, P, U64, U64, U32, U64, U64, C.POINTER(P)]),
    "ordvec_server_info": (ST, [P, P, P, P]),
    "ordvec_server_submit": (ST, [P, P, U64, U64, P, P, P]),
    "ordvec_server_wait": (ST, [P, U64, P, P]),
    "ordvec_server_open_loop": (ST, [P, P, U64, C.c_double, C.c_double, U32, U32, U32, U32, U64, P]),
    "ordvec_server_stop": (None, [P]),
    "ordvec_rankquant_search": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_rankquant_swap_remove": (ST, [P, U64, P]),
    "ordvec_sign_bitmap_swap_remove": (ST, [P, U64, P]),
    "ordvec_bitmap_swap_remove": (ST, [P, U64, P]),
    "ordvec_rank_new": (ST, [U64, C.POINTER(P)]),
    "ordvec_rank_add": (ST, [P, P, U64, U64]),
    "ordvec_rank_info": (ST, [P, P, P]),
    "ordvec_rank_search": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_rank_search_asymmetric": (ST, [P, P, U64, U64, P, P, U64]),
    "ordvec_rank_ranks": (ST, [P, P, U64, U64]),
    "ordvec_rank_swap_remove": (ST, [P, U64, P]),
    "ordvec_rank_write": (ST, [P, C.c_char_p]),
    "ordvec_rank_load": (ST, [C.c_char_p, C.POINTER(P)]),
    "ordvec_rank_free": (None, [P]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load and declare the native library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own HIP runtime (same soname as ROCm's).  Loading torch
    # first makes this library bind to that runtime, so the process holds
    # exactly one HIP runtime; two runtimes in one process cannot share the GPU.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    if not os.path.exists(path):
        raise ImportError(
            f"ordvec_amd native library not found at {path}; build it with "
            "`python ordvec_amd/_build.py` (hipcc, gfx950)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return _lib if _lib is not None else load_library()


def last_error() -> str:
    msg = lib().ordvec_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(status: int) -> None:
    if status == OK:
        return
    msg = last_error()
    name = lib().ordvec_status_name(status).decode()
    text = f"{name}: {msg}"
    if status in (NULL_POINTER, BAD_ARGUMENT, BAD_STRUCT_SIZE, DIM_MISMATCH, NONFINITE_QUERY,
                  ROW_ID_OUT_OF_RANGE, BUFFER_TOO_SMALL):
        raise OrdvecValueError(status, text)
    if status in (UNSUPPORTED_FORMAT, CORRUPT_INDEX, IO):
        raise OrdvecIOError(status, text)
    if status == UNSUPPORTED_OPERATION:
        raise OrdvecUnsupported(status, text)
    raise OrdvecError(status, text)


KERNEL_FAMILIES = {"sign_scan": 0, "stage1_select": 1, "rerank": 2, "full_scan": 3, "fastscan": 4,
                   "encode": 5, "prepare": 6, "stage1_sample": 7, "backward": 8}


ENGINE_OPTIONS = {"one_query_path": 1, "mfma_scan": 2, "batch_tail": 3, "side_prep": 4}


def engine_option(name: str, value=None) -> bool:
    """Switch between equivalent execution paths (ordvec_engine_option):
    value None only reads; returns the setting before the call."""
    prev = C.c_int64()
    check(lib().ordvec_engine_option(ENGINE_OPTIONS[name], -1 if value is None else int(bool(value)),
                                     C.byref(prev)))
    return bool(prev.value)


def kernel_timing_enable(on: bool = True, families=None) -> None:
    """Reset the counters and record events around the launches of every
    kernel family (``families`` None) or of the named families only."""
    if families is None:
        check(lib().ordvec_kernel_timing_enable(1 if on else 0))
    else:
        mask = 0
        for f in families:
            mask |= 1 << KERNEL_FAMILIES[f]
        check(lib().ordvec_kernel_timing_enable_families(mask if on else 0))


def kernel_timing_read(family: str):
    """(launches, total GPU ms) of a kernel family since enable."""
    n = C.c_uint64()
    ms = C.c_double()
    check(lib().ordvec_kernel_timing_read(KERNEL_FAMILIES[family], C.byref(n), C.byref(ms)))
    return n.value, ms.value
