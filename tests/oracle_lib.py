"""Loader for the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_char_p, c_double, c_int, c_uint64, c_void_p

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "issl_ref_harness")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.or_encode.restype = c_uint64
        L.or_encode.argtypes = [c_char_p]
        L.or_mit_local.restype = c_double
        L.or_mit_local.argtypes = [c_void_p, c_uint64]
        L.or_mit_fold.restype = c_double
        L.or_mit_fold.argtypes = [c_uint64]
        L.or_mit_table.restype = c_uint64
        L.or_mit_table.argtypes = [c_void_p, c_void_p]
        L.or_cfd_tables.restype = None
        L.or_cfd_tables.argtypes = [c_void_p, c_void_p]
        L.or_score.restype = c_int
        L.or_score.argtypes = [c_void_p, c_uint64, c_uint64, c_void_p, c_uint64, c_void_p,
                               c_void_p, c_void_p, c_uint64, c_int, c_double, c_int,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p]
        L.or_candidates_brute.restype = c_uint64
        L.or_candidates_brute.argtypes = [c_void_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                          c_uint64, c_void_p, c_uint64]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def mit_table():
    k = np.zeros(6195, dtype=np.uint64)
    v = np.zeros(6195, dtype=np.float64)
    lib().or_mit_table(_p(k), _p(v))
    return k, v


def cfd_tables():
    p = np.zeros(320)
    a = np.zeros(16)
    lib().or_cfd_tables(_p(p), _p(a))
    return p, a


def score(ix, guides, method: int, threshold: float, threads: int = 0, distinct: bool = False):
    """Oracle scores; returns a dict of per-guide arrays (and the distinct bitmap)."""
    g = np.ascontiguousarray(guides, dtype=np.uint64)
    n = g.shape[0]
    out = {k: np.zeros(n) for k in ("mit", "cfd", "mit_sum", "cfd_sum")}
    out["occ_count"] = np.zeros(n, dtype=np.uint64)
    out["site_count"] = np.zeros(n, dtype=np.uint64)
    out["pass"] = np.zeros(n, dtype=np.uint8)
    bm = np.zeros(ix.n_sites // 64 + 1, dtype=np.uint64) if distinct else None
    rc = lib().or_score(_p(ix.sigs), ix.n_sites, ix.seq_len, _p(ix.masks), ix.n_slices,
                        _p(ix.list_sizes), _p(ix.entries), _p(g), n, method, threshold, threads,
                        _p(out["mit"]), _p(out["cfd"]), _p(out["occ_count"]),
                        _p(out["site_count"]), _p(out["pass"]), _p(out["mit_sum"]),
                        _p(out["cfd_sum"]), _p(bm))
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    if distinct:
        out["distinct"] = bm
    return out


def candidates_brute(ix, guide: int, cap: int = 1 << 20) -> np.ndarray:
    ids = np.zeros(cap, dtype=np.uint32)
    n = lib().or_candidates_brute(_p(ix.sigs), ix.n_sites, ix.seq_len, _p(ix.masks),
                                  ix.n_slices, int(guide), _p(ids), cap)
    if n > cap:
        raise ValueError("capacity")
    return ids[:n]
