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
                               c_void_p, c_void_p, c_void_p]
        L.or_extract.restype = c_uint64
        L.or_extract.argtypes = [c_char_p, c_uint64, c_void_p, c_uint64]
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


def score(ix, guides, method: int, threshold: float, threads: int = 0, distinct: bool = False,
          margin: bool = False):
    """Oracle scores; returns a dict of per-guide arrays (and the distinct bitmap,
    and the exit margin of or_score when asked)."""
    g = np.ascontiguousarray(guides, dtype=np.uint64)
    n = g.shape[0]
    out = {k: np.zeros(n) for k in ("mit", "cfd", "mit_sum", "cfd_sum")}
    out["occ_count"] = np.zeros(n, dtype=np.uint64)
    out["site_count"] = np.zeros(n, dtype=np.uint64)
    out["pass"] = np.zeros(n, dtype=np.uint8)
    bm = np.zeros(ix.n_sites // 64 + 1, dtype=np.uint64) if distinct else None
    mg = np.zeros(n) if margin else None
    rc = lib().or_score(_p(ix.sigs), ix.n_sites, ix.seq_len, _p(ix.masks), ix.n_slices,
                        _p(ix.list_sizes), _p(ix.entries), _p(g), n, method, threshold, threads,
                        _p(out["mit"]), _p(out["cfd"]), _p(out["occ_count"]),
                        _p(out["site_count"]), _p(out["pass"]), _p(out["mit_sum"]),
                        _p(out["cfd_sum"]), _p(bm), _p(mg))
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    if distinct:
        out["distinct"] = bm
    if margin:
        out["margin"] = mg
    return out


def candidates_brute(ix, guide: int, cap: int = 1 << 20) -> np.ndarray:
    ids = np.zeros(cap, dtype=np.uint32)
    n = lib().or_candidates_brute(_p(ix.sigs), ix.n_sites, ix.seq_len, _p(ix.masks),
                                  ix.n_slices, int(guide), _p(ids), cap)
    if n > cap:
        raise ValueError("capacity")
    return ids[:n]


# ---------------------------------------------------------------- extraction

_WS = b" \t\n\v\f\r"


def split_input(text: bytes) -> list:
    """Sequences of one input file as ExtractOfftargets.cpp:76-109 splits
    them: FASTA when the first line starts with '>' (every later line
    trimmed and upper-cased; a '>' line starts the next sequence, other
    lines are appended), otherwise every line (trimmed, upper-cased) is a
    sequence of its own."""
    lines = text.split(b"\n")
    if text.endswith(b"\n"):
        lines = lines[:-1]   # getline yields no empty last line
    if not lines:
        return [b""]         # the reference's do-while writes one empty file
    out = []
    if lines[0][:1] == b">":
        cur = []
        for ln in lines[1:]:
            ln = ln.strip(_WS).upper()
            if ln[:1] == b">":
                out.append(b"".join(cur))
                cur = []
            else:
                cur.append(ln)
        out.append(b"".join(cur))
    else:
        out = [ln.strip(_WS).upper() for ln in lines]
    return out


def extract(seq: bytes) -> list:
    """Sites of one sequence (or_extract): forward ones in position order,
    then reverse ones."""
    n = lib().or_extract(seq, len(seq), None, 0)
    buf = ctypes.create_string_buffer(max(1, 20 * n))
    lib().or_extract(seq, len(seq), ctypes.cast(buf, c_void_p), n)
    raw = buf.raw[: 20 * n].decode()
    return [raw[20 * i: 20 * i + 20] for i in range(n)]


def extract_output(texts) -> bytes:
    """The reference's output file for these input files: every site of
    every sequence, sorted, one per line (duplicates kept)."""
    sites = []
    for t in texts:
        for seq in split_input(t):
            sites.extend(extract(seq))
    sites.sort()
    return "".join(x + "\n" for x in sites).encode()
