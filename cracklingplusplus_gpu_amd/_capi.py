"""ctypes binding of libissl_gpu.so (include/issl_gpu.h, issl_index.h, issl_synth.h).

The library is the product: every call here goes straight to the C ABI.  If
the shared library has not been built, loading fails loudly; there is no
fallback path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char, c_char_p, c_double, c_int, c_int32,
                    c_uint8, c_uint32, c_uint64, c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libissl_gpu.so")

ISSL_OK = 0
ISSL_ERR_ARG = -1
ISSL_ERR_IO = -2
ISSL_ERR_FORMAT = -3
ISSL_ERR_DEVICE = -4
ISSL_ERR_NOMEM = -5
ISSL_ERR_CAPACITY = -6
ISSL_ERR_TIMEOUT = -7
ISSL_ERR_STATE = -8

METHOD_MIT, METHOD_CFD, METHOD_MIT_AND_CFD, METHOD_MIT_OR_CFD, METHOD_AVG_MIT_CFD = range(5)
FORM_PLAIN, FORM_BATCHED, FORM_ORDERED = 0, 1, 2
PREC_FP64, PREC_FP32C = 0, 1
FLAG_DISTINCT = 1
FLAG_GRAPH = 2


class issl_params(Structure):
    _fields_ = [("max_dist", c_int32), ("method", c_int32), ("threshold", c_double),
                ("form", c_int32), ("precision", c_int32), ("flags", c_uint32),
                ("chunk", c_int32)]


class issl_results(Structure):
    _fields_ = [("mit", c_void_p), ("cfd", c_void_p), ("occ_count", c_void_p),
                ("site_count", c_void_p), ("pass_", c_void_p), ("mit_sum", c_void_p),
                ("cfd_sum", c_void_p)]


class issl_index_info(Structure):
    _fields_ = [("n_sites", c_uint64), ("seq_len", c_uint64), ("n_slices", c_uint64),
                ("n_lists", c_uint64), ("device_bytes", c_uint64), ("device", c_int32),
                ("max_width", c_int32)]


class issl_tuning(Structure):
    _fields_ = [("chunk", c_int32), ("block", c_int32), ("depth", c_int32),
                ("grid_per_cu", c_int32), ("direct_depth", c_int32),
                ("heavy_kpost", c_int32),
                ("us_per_guide", c_double),
                ("ord_depth", c_int32), ("ord_grid", c_int32), ("ord_split", c_int32),
                ("ovl_grid", c_int32)]


HEAVY_NEVER = 0x7FFFFFFF
SPLIT_NEVER = 0x7FFFFFFF


class issl_synth_cfg(Structure):
    _fields_ = [("n_sites", c_uint64), ("gc", c_double), ("repeat_frac", c_double),
                ("n_repeat_seeds", c_uint64), ("repeat_sub_p", c_double),
                ("occ_p_unique", c_double), ("occ_p_repeat", c_double),
                ("n_slices", c_uint64), ("slice_width", c_uint64),
                ("contiguous_masks", c_int32), ("guide_batches", c_int32),
                ("n_guides_site", c_uint64), ("n_guides_unique", c_uint64),
                ("n_guides_planted", c_uint64), ("n_guides_repeat", c_uint64),
                ("n_guides_random", c_uint64), ("seed", c_uint64), ("guide_sub_p", c_double)]


class issl_server_config(Structure):
    _fields_ = [("device", c_int32), ("slots", c_int32), ("max_batch", c_uint64),
                ("hbm_reserve", c_uint64)]


class issl_server_stats(Structure):
    _fields_ = [("requests_done", c_uint64), ("requests_failed", c_uint64),
                ("guides_done", c_uint64), ("launches", c_uint64), ("queued", c_uint64),
                ("peak_inflight", c_uint64), ("resident_bytes", c_uint64), ("hbm_free", c_uint64), ("hbm_total", c_uint64),
                ("p50_ms", c_double), ("p99_ms", c_double), ("max_ms", c_double),
                ("mean_ms", c_double), ("mean_batch_guides", c_double)]


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    # issl_gpu.h
    ("issl_index_open", c_int, [c_char_p, c_int, POINTER(c_void_p)]),
    ("issl_index_from_arrays", c_int, [c_uint64, c_uint64, c_uint64, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_int, POINTER(c_void_p)]),
    ("issl_index_close", c_int, [c_void_p]),
    ("issl_index_open_streamed", c_int, [c_char_p, c_int, c_uint64, POINTER(c_void_p)]),
    ("issl_index_from_arrays_streamed", c_int, [c_uint64, c_uint64, c_uint64, c_void_p, c_void_p,
                                                c_void_p, c_void_p, c_int, c_uint64,
                                                POINTER(c_void_p)]),
    ("issl_index_stream_stats", c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64),
                                        POINTER(c_uint64)]),
    ("issl_index_build", c_int, [c_uint64, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p,
                                 c_int, POINTER(c_void_p)]),
    ("issl_index_export", c_int, [c_void_p, c_void_p, c_void_p]),
    ("issl_index_get_info", c_int, [c_void_p, POINTER(issl_index_info)]),
    ("issl_index_validate", c_int, [c_void_p, POINTER(c_uint64)]),
    ("issl_index_set_residual", c_int, [c_void_p, c_int32]),
    ("issl_index_bind_stream", c_int, [c_void_p, c_void_p]),
    ("issl_index_unbind_stream", c_int, [c_void_p, c_void_p]),
    ("issl_index_get_residual", c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32),
                                        POINTER(c_uint32)]),
    ("issl_distinct_count", c_int, [c_void_p, POINTER(c_uint64)]),
    ("issl_distinct_reset", c_int, [c_void_p]),
    ("issl_encode", c_int, [c_char_p, c_uint64, c_uint64, c_void_p]),
    ("issl_score", c_int, [c_void_p, c_void_p, c_uint64, POINTER(issl_params),
                           POINTER(issl_results)]),
    ("issl_score_device", c_int, [c_void_p, c_void_p, c_uint64, POINTER(issl_params),
                                  POINTER(issl_results), c_void_p]),
    ("issl_score_device_noalloc", c_int, [c_void_p, c_void_p, c_uint64, POINTER(issl_params),
                                          POINTER(issl_results), c_void_p]),
    ("issl_candidates", c_int, [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint64]),
    ("issl_autotune", c_int, [c_void_p, c_void_p, c_uint64, c_char_p, POINTER(issl_tuning)]),
    ("issl_autotune_precision", c_int, [c_void_p, c_void_p, c_uint64, c_int32, c_char_p,
                                        POINTER(issl_tuning)]),
    ("issl_get_tuning", c_int, [c_void_p, POINTER(issl_tuning)]),
    ("issl_set_tuning", c_int, [c_void_p, POINTER(issl_tuning)]),
    ("issl_timing_enable", c_int, [c_void_p, c_int]),
    ("issl_timing_read", c_int, [c_void_p, POINTER(c_double), POINTER(c_uint64)]),
    ("issl_last_error", c_char_p, []),
    ("issl_device_count", c_int, []),
    ("issl_version", c_char_p, []),
    # issl_index.h
    ("issl_list_count", c_uint64, [c_void_p, c_uint64, c_uint64]),
    ("issl_build_lists", c_int, [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                 c_void_p, c_void_p, c_int]),
    ("issl_write_index", c_int, [c_char_p, c_uint64, c_uint64, c_uint64, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    ("issl_read_header", c_int, [c_char_p, POINTER(c_uint64), POINTER(c_uint64),
                                 POINTER(c_uint64), POINTER(c_uint64)]),
    ("issl_read_index", c_int, [c_char_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("issl_parse_masks", c_int, [c_char_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                 POINTER(c_uint64)]),
    ("issl_count_sorted_sites", c_int, [c_char_p, c_uint64, c_uint64, c_uint64, c_void_p,
                                        c_void_p, c_uint64, POINTER(c_uint64)]),
    ("issl_score_tables", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    # issl_server.h
    ("issl_extractor_create", c_int, [c_int, c_uint64, POINTER(c_void_p)]),
    ("issl_extractor_destroy", c_int, [c_void_p]),
    ("issl_extractor_add_sequence", c_int, [c_void_p, c_void_p, c_uint64]),
    ("issl_extractor_add_path", c_int, [c_void_p, c_char_p]),
    ("issl_extractor_finish", c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64)]),
    ("issl_extractor_sites", c_int, [c_void_p, c_void_p, c_void_p]),
    ("issl_extractor_write", c_int, [c_void_p, c_char_p]),
    ("issl_extractor_build_index", c_int, [c_void_p, c_void_p, c_uint64, POINTER(c_void_p)]),
    ("issl_extractor_stats", c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64)]),
    ("issl_server_create", c_int, [POINTER(issl_server_config), POINTER(c_void_p)]),
    ("issl_server_destroy", c_int, [c_void_p]),
    ("issl_server_add_index", c_int, [c_void_p, c_char_p, POINTER(c_int32)]),
    ("issl_server_add_index_arrays", c_int, [c_void_p, c_uint64, c_uint64, c_uint64, c_void_p,
                                             c_void_p, c_void_p, c_void_p, POINTER(c_int32)]),
    ("issl_server_remove_index", c_int, [c_void_p, c_int32]),
    ("issl_server_submit", c_int, [c_void_p, c_int32, c_void_p, c_uint64, POINTER(issl_params),
                                   POINTER(issl_results), POINTER(c_uint64)]),
    ("issl_server_wait", c_int, [c_void_p, c_uint64, c_double, POINTER(c_int32)]),
    ("issl_server_poll", c_int, [c_void_p, c_uint64, POINTER(c_int32), POINTER(c_int32)]),
    ("issl_server_get_stats", c_int, [c_void_p, c_int, POINTER(issl_server_stats)]),
    ("issl_device_memory", c_int, [c_int, POINTER(c_uint64), POINTER(c_uint64)]),
    ("issl_index_bytes_needed", c_uint64, [c_uint64, c_uint64, c_uint64]),
    # issl_synth.h
    ("issl_synth", c_int, [POINTER(issl_synth_cfg), c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_int]),
]

_lib = None


def lib():
    """Load libissl_gpu.so once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run __graft_entry__.build() "
                "(make -C cracklingplusplus_gpu_amd/csrc)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class IsslError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc: int) -> None:
    if rc != ISSL_OK:
        msg = lib().issl_last_error()
        raise IsslError(rc, msg.decode() if msg else "")
