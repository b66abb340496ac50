"""MI355X ISSL off-target scorer (MIT + CFD specificity).

Python face of libissl_gpu.so, used by the tests and the benchmark.  The
scorer itself is HIP/C++ behind the C ABI in include/issl_gpu.h; this module
only moves numpy arrays across it.

Drop-in surface of the reference:
  * ISSLScoringModule::run  (lib/ISSLScoringModule.cpp:52-466)  -> Index.score
  * ISSLScoreOfftargets CLI (ISSLScoreOfftargets.cpp:36-647)    -> bin/ISSLScoreOfftargets
  * ISSLCreateIndex          (ISSLCreateIndex.cpp:284-493)       -> build_lists / write_index
"""
from __future__ import annotations

import ctypes
import os
from ctypes import byref, c_uint64, c_void_p
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _capi
from ._capi import (FLAG_DISTINCT, FLAG_GRAPH, FORM_BATCHED, FORM_ORDERED, FORM_PLAIN, METHOD_AVG_MIT_CFD, METHOD_CFD,
                    METHOD_MIT, METHOD_MIT_AND_CFD, METHOD_MIT_OR_CFD, PREC_FP32C, PREC_FP64,
                    SPLIT_NEVER, IsslError, check, lib)

__all__ = [
    "Index", "IndexArrays", "Results", "IsslError", "encode", "decode", "build_lists",
    "write_index", "read_index", "parse_masks", "count_sorted_sites", "score_tables",
    "synth", "device_count", "METHOD_MIT", "METHOD_CFD", "METHOD_MIT_AND_CFD",
    "METHOD_MIT_OR_CFD", "METHOD_AVG_MIT_CFD", "FORM_PLAIN", "FORM_BATCHED", "FORM_ORDERED", "PREC_FP64",
    "PREC_FP32C", "FLAG_DISTINCT", "FLAG_GRAPH", "METHOD_NAMES", "CLI_METHODS", "Server", "Request",
    "device_memory", "Extractor",
]

# include/util.hpp:20 and lib/util.cpp:13-19
METHOD_NAMES = {"mit": METHOD_MIT, "cfd": METHOD_CFD, "mitAndCfd": METHOD_MIT_AND_CFD,
                "mitOrCfd": METHOD_MIT_OR_CFD, "avgMitCfd": METHOD_AVG_MIT_CFD}
# ISSLScoreOfftargets.cpp:69-91
CLI_METHODS = {"and": METHOD_MIT_AND_CFD, "or": METHOD_MIT_OR_CFD, "avg": METHOD_AVG_MIT_CFD,
               "mit": METHOD_MIT, "cfd": METHOD_CFD}


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(c_void_p)


def device_count() -> int:
    return int(lib().issl_device_count())


# ------------------------------------------------------------------ encoding

_CODE = np.zeros(256, dtype=np.uint64)
_CODE[ord("C")] = 1
_CODE[ord("G")] = 2
_CODE[ord("T")] = 3


def encode(seqs) -> np.ndarray:
    """20-mers -> signatures through issl_encode (ISSLScoringModule.cpp:469-476)."""
    seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    n = len(seqs)
    out = np.zeros(n, dtype=np.uint64)
    if n == 0:
        return out
    if any(len(s) < 20 for s in seqs):
        raise ValueError("guides must have at least 20 characters")
    buf = b"".join(s[:20] for s in seqs)
    check(lib().issl_encode(buf, n, 20, _ptr(out)))
    return out


def decode(sigs) -> list:
    """Signatures -> 20-mers (ISSLScoringModule.cpp:478-485)."""
    out = []
    for s in np.asarray(sigs, dtype=np.uint64).tolist():
        out.append("".join("ACGT"[(s >> (2 * j)) & 3] for j in range(20)))
    return out


# ------------------------------------------------------------------ index data

@dataclass
class IndexArrays:
    """Host arrays of an index in the .issl layout (ISSLCreateIndex.cpp:428-484)."""
    sigs: np.ndarray          # uint64[N]
    masks: np.ndarray         # uint64[S]
    list_sizes: np.ndarray    # uint64[sum 4^w]
    entries: np.ndarray       # uint64[S*N]  occ<<32 | id
    seq_len: int = 20

    @property
    def n_sites(self) -> int:
        return int(self.sigs.shape[0])

    @property
    def n_slices(self) -> int:
        return int(self.masks.shape[0])

    def occurrences(self) -> np.ndarray:
        """occ per id, read from slice 0's entries."""
        occ = np.zeros(self.n_sites, dtype=np.uint32)
        if self.n_slices:
            e = self.entries[: self.n_sites]
            occ[(e & np.uint64(0xFFFFFFFF)).astype(np.int64)] = (e >> np.uint64(32)).astype(np.uint32)
        return occ


def list_count(masks, seq_len: int = 20) -> int:
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    return int(lib().issl_list_count(_ptr(m), m.shape[0], seq_len))


def build_lists(sigs, occ, masks, seq_len: int = 20, threads: int = 0) -> IndexArrays:
    """Slice lists of distinct sites (ISSLCreateIndex.cpp:450-484)."""
    sigs = np.ascontiguousarray(sigs, dtype=np.uint64)
    occ = np.ascontiguousarray(occ, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint64)
    n, s = sigs.shape[0], masks.shape[0]
    nl = list_count(masks, seq_len) if s else 0
    if s and nl == 0:
        raise ValueError("slice wider than 16 positions")
    sizes = np.zeros(max(nl, 1), dtype=np.uint64)[:nl]
    entries = np.zeros(max(n * s, 1), dtype=np.uint64)[: n * s]
    check(lib().issl_build_lists(_ptr(sigs), _ptr(occ), n, seq_len, _ptr(masks), s,
                                 _ptr(sizes), _ptr(entries), threads))
    return IndexArrays(sigs, masks, sizes, entries, seq_len)


def write_index(path: str, ix: IndexArrays) -> None:
    check(lib().issl_write_index(path.encode(), ix.n_sites, ix.seq_len, ix.n_slices,
                                 _ptr(ix.sigs), _ptr(ix.masks), _ptr(ix.list_sizes),
                                 _ptr(ix.entries)))


def read_index(path: str) -> IndexArrays:
    n, L, S, nl = c_uint64(), c_uint64(), c_uint64(), c_uint64()
    check(lib().issl_read_header(path.encode(), byref(n), byref(L), byref(S), byref(nl)))
    sigs = np.zeros(n.value, dtype=np.uint64)
    masks = np.zeros(S.value, dtype=np.uint64)
    sizes = np.zeros(nl.value, dtype=np.uint64)
    entries = np.zeros(n.value * S.value, dtype=np.uint64)
    check(lib().issl_read_index(path.encode(), _ptr(sigs), _ptr(masks), _ptr(sizes),
                                _ptr(entries)))
    return IndexArrays(sigs, masks, sizes, entries, int(L.value))


def parse_masks(text: str, seq_len: int = 20) -> np.ndarray:
    b = text.encode()
    n = c_uint64()
    check(lib().issl_parse_masks(b, len(b), seq_len, None, 0, byref(n)))
    out = np.zeros(n.value, dtype=np.uint64)
    check(lib().issl_parse_masks(b, len(b), seq_len, _ptr(out), n.value, byref(n)))
    return out


def count_sorted_sites(text: bytes, seq_len: int = 20, line_len: int = 21):
    n = c_uint64()
    check(lib().issl_count_sorted_sites(text, len(text), seq_len, line_len, None, None, 0,
                                        byref(n)))
    sigs = np.zeros(n.value, dtype=np.uint64)
    occ = np.zeros(n.value, dtype=np.uint32)
    check(lib().issl_count_sorted_sites(text, len(text), seq_len, line_len, _ptr(sigs),
                                        _ptr(occ), n.value, byref(n)))
    return sigs, occ


def score_tables():
    """(mit_keys[6195], mit_vals[6195], cfd_pos[320], cfd_pam[16]) as the product holds them."""
    k = np.zeros(6195, dtype=np.uint64)
    v = np.zeros(6195, dtype=np.float64)
    p = np.zeros(320, dtype=np.float64)
    a = np.zeros(16, dtype=np.float64)
    check(lib().issl_score_tables(_ptr(k), _ptr(v), _ptr(p), _ptr(a)))
    return k, v, p, a


def synth(cfg: dict, threads: int = 0):
    """Generate a seeded workload: (sigs, occ, masks, guides, kinds).  With
    guide_batches = B > 1, guides holds B disjoint batches of the stated mix,
    one after another (batch 0 is the single-batch workload)."""
    c = _capi.issl_synth_cfg()
    for key, val in cfg.items():
        setattr(c, key, val)
    q = (c.n_guides_site + c.n_guides_unique + c.n_guides_planted + c.n_guides_repeat
         + c.n_guides_random) * max(1, c.guide_batches)
    sigs = np.zeros(c.n_sites, dtype=np.uint64)
    occ = np.zeros(c.n_sites, dtype=np.uint32)
    masks = np.zeros(c.n_slices, dtype=np.uint64)
    guides = np.zeros(q, dtype=np.uint64)
    kinds = np.zeros(q, dtype=np.uint8)
    check(lib().issl_synth(byref(c), _ptr(sigs), _ptr(occ), _ptr(masks), _ptr(guides),
                           _ptr(kinds), threads))
    return sigs, occ, masks, guides, kinds


# ------------------------------------------------------------------ scoring

@dataclass
class Results:
    mit: np.ndarray
    cfd: np.ndarray
    occ_count: np.ndarray
    site_count: np.ndarray
    passed: np.ndarray      # bytes b'1' / b'0' as uint8
    mit_sum: np.ndarray
    cfd_sum: np.ndarray


class Index:
    """A .issl index resident on one GPU (issl_index_open / issl_index_from_arrays)."""

    def __init__(self, handle: int, keep=None):
        self._h = c_void_p(handle)
        self._keep = keep  # host arrays a streamed index reads from

    @classmethod
    def open(cls, path: str, device: int = 0) -> "Index":
        h = c_void_p()
        check(lib().issl_index_open(path.encode(), device, byref(h)))
        return cls(h.value)

    @classmethod
    def open_streamed(cls, path: str, device: int = 0, slot_bytes: int = 0) -> "Index":
        """Memory-map a .issl file and stream its slices through `slot_bytes`
        of device slice buffers (issl_index_open_streamed; 0: as many as fit)."""
        h = c_void_p()
        check(lib().issl_index_open_streamed(path.encode(), device, slot_bytes, byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays_streamed(cls, ix: IndexArrays, device: int = 0,
                             slot_bytes: int = 0) -> "Index":
        """Streamed index over host arrays (issl_index_from_arrays_streamed); the
        arrays are kept referenced until the index is closed."""
        for a in (ix.sigs, ix.masks, ix.list_sizes, ix.entries):
            if a.dtype != np.uint64 or not a.flags.c_contiguous:
                raise ValueError("streamed index arrays must be C-contiguous uint64")
        h = c_void_p()
        check(lib().issl_index_from_arrays_streamed(
            ix.n_sites, ix.seq_len, ix.n_slices, _ptr(ix.sigs), _ptr(ix.masks),
            _ptr(ix.list_sizes), _ptr(ix.entries), device, slot_bytes, byref(h)))
        return cls(h.value, keep=ix)

    def stream_stats(self):
        """(slice buffers, bytes copied into them, windows scored); zeros for a
        resident index."""
        a, b, c = c_uint64(), c_uint64(), c_uint64()
        check(lib().issl_index_stream_stats(self._h, byref(a), byref(b), byref(c)))
        return a.value, b.value, c.value

    @classmethod
    def from_arrays(cls, ix: IndexArrays, device: int = 0) -> "Index":
        h = c_void_p()
        check(lib().issl_index_from_arrays(ix.n_sites, ix.seq_len, ix.n_slices, _ptr(ix.sigs),
                                           _ptr(ix.masks), _ptr(ix.list_sizes),
                                           _ptr(ix.entries), device, byref(h)))
        return cls(h.value)

    @classmethod
    def build(cls, sigs, occ, masks, seq_len: int = 20, device: int = 0) -> "Index":
        """Build the slice lists on the device from distinct sites (issl_index_build)."""
        sigs = np.ascontiguousarray(sigs, dtype=np.uint64)
        occ = np.ascontiguousarray(occ, dtype=np.uint32)
        masks = np.ascontiguousarray(masks, dtype=np.uint64)
        h = c_void_p()
        check(lib().issl_index_build(sigs.shape[0], seq_len, masks.shape[0], _ptr(sigs),
                                     _ptr(occ), _ptr(masks), device, byref(h)))
        return cls(h.value)

    def export(self, entries: bool = True):
        """(list_sizes, entries) of the resident index in the .issl layout
        (entries None when not requested)."""
        info = self.info()
        sizes = np.zeros(max(info.n_lists, 1), dtype=np.uint64)[: info.n_lists]
        ent = None
        if entries:
            ent = np.zeros(max(info.n_sites * info.n_slices, 1), dtype=np.uint64)[
                : info.n_sites * info.n_slices]
        check(lib().issl_index_export(self._h, _ptr(sizes), _ptr(ent)))
        return sizes, ent

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h is not None and self._h.value:
            check(lib().issl_index_close(self._h))
        self._h = None
        self._keep = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> _capi.issl_index_info:
        out = _capi.issl_index_info()
        check(lib().issl_index_get_info(self._h, byref(out)))
        return out

    def set_residual(self, use: bool) -> None:
        """Stream the list-residual words (4 B per posting) in the BATCHED tile
        kernel when the index has them, or the signature-inline words (8 B, a
        new index's choice; the autotuner picks the faster).  Results are the
        same bits either way."""
        check(lib().issl_index_set_residual(self._h, 1 if use else 0))

    def residual(self) -> tuple:
        """(built, in use, occurrence bits per residual word)."""
        b, u, o = _capi.c_int32(), _capi.c_int32(), _capi.c_uint32()
        check(lib().issl_index_get_residual(self._h, byref(b), byref(u), byref(o)))
        return bool(b.value), bool(u.value), int(o.value)

    def validate(self) -> int:
        bad = c_uint64()
        check(lib().issl_index_validate(self._h, byref(bad)))
        return int(bad.value)

    def distinct_count(self) -> int:
        c = c_uint64()
        check(lib().issl_distinct_count(self._h, byref(c)))
        return int(c.value)

    def distinct_reset(self) -> None:
        check(lib().issl_distinct_reset(self._h))

    @staticmethod
    def params(method: int = METHOD_MIT_AND_CFD, threshold: float = 0.0,
               form: int = FORM_PLAIN, precision: int = PREC_FP64, flags: int = 0,
               chunk: int = 0, max_dist: int = 4) -> _capi.issl_params:
        return _capi.issl_params(max_dist, method, threshold, form, precision, flags, chunk)

    @staticmethod
    def results(n: int) -> Results:
        """Host result buffers for n guides (score's `out`)."""
        return Results(np.zeros(n), np.zeros(n), np.zeros(n, np.uint64), np.zeros(n, np.uint64),
                       np.zeros(n, np.uint8), np.zeros(n), np.zeros(n))

    def score(self, guides, method: int = METHOD_MIT_AND_CFD, threshold: float = 0.0,
              form: int = FORM_PLAIN, precision: int = PREC_FP64, flags: int = 0,
              chunk: int = 0, max_dist: int = 4, out: Optional[Results] = None) -> Results:
        """Blocking scoring through host buffers (issl_score).  `out`: buffers
        from results(n) to fill instead of new ones."""
        g = np.ascontiguousarray(guides, dtype=np.uint64)
        n = g.shape[0]
        r = out if out is not None else self.results(n)
        for a, dt in ((r.mit, np.float64), (r.cfd, np.float64), (r.occ_count, np.uint64),
                      (r.site_count, np.uint64), (r.passed, np.uint8), (r.mit_sum, np.float64),
                      (r.cfd_sum, np.float64)):
            if a.shape != (n,) or a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError("out buffers must be results(n) for this call's n guides")
        res = _capi.issl_results(_ptr(r.mit), _ptr(r.cfd), _ptr(r.occ_count),
                                 _ptr(r.site_count), _ptr(r.passed), _ptr(r.mit_sum),
                                 _ptr(r.cfd_sum))
        p = self.params(method, threshold, form, precision, flags, chunk, max_dist)
        check(lib().issl_score(self._h, _ptr(g), n, byref(p), byref(res)))
        return r

    def candidates(self, guides):
        """CSR candidate sets (offsets[n+1], ids) in the reference's visit order."""
        g = np.ascontiguousarray(guides, dtype=np.uint64)
        n = g.shape[0]
        off = np.zeros(n + 1, dtype=np.uint64)
        rc = lib().issl_candidates(self._h, _ptr(g), n, _ptr(off), None, 0)
        if rc not in (_capi.ISSL_OK, _capi.ISSL_ERR_CAPACITY):
            check(rc)
        total = int(off[-1])
        ids = np.zeros(max(total, 1), dtype=np.uint32)
        check(lib().issl_candidates(self._h, _ptr(g), n, _ptr(off), _ptr(ids), total))
        return off, ids[:total]

    def autotune(self, sample, cache_path: Optional[str] = None,
                 precision: int = PREC_FP64) -> _capi.issl_tuning:
        """Time the BATCHED launch shapes and site-word layouts on `sample` at
        `precision` and keep the fastest (issl_autotune_precision)."""
        g = np.ascontiguousarray(sample, dtype=np.uint64)
        t = _capi.issl_tuning()
        check(lib().issl_autotune_precision(self._h, _ptr(g), g.shape[0], precision,
                                            cache_path.encode() if cache_path else None,
                                            byref(t)))
        return t

    def tuning(self) -> _capi.issl_tuning:
        t = _capi.issl_tuning()
        check(lib().issl_get_tuning(self._h, byref(t)))
        return t

    def timing_enable(self, enable=True, repeat: int = 1) -> None:
        """Time the dominant kernel of each scoring launch; repeat > 1 dispatches it
        that many times back to back between one event pair (results unchanged)."""
        check(lib().issl_timing_enable(self._h, max(1, int(repeat)) if enable else 0))

    def timing_read(self):
        """(summed dominant-kernel ms, dispatches timed) since timing_enable."""
        ms, n = ctypes.c_double(), c_uint64()
        check(lib().issl_timing_read(self._h, byref(ms), byref(n)))
        return ms.value, int(n.value)

    def score_device(self, d_guides, n: int, d_out: dict, method: int = METHOD_MIT_AND_CFD,
                     threshold: float = 0.0, form: int = FORM_PLAIN,
                     precision: int = PREC_FP64, flags: int = 0, chunk: int = 0,
                     stream=None) -> None:
        """Asynchronous scoring of device-resident guides into device buffers.

        d_guides / d_out values are hiprt.DeviceArray (or None per output)."""
        def dp(x):
            return None if x is None else x.ptr
        res = _capi.issl_results(dp(d_out.get("mit")), dp(d_out.get("cfd")),
                                 dp(d_out.get("occ_count")), dp(d_out.get("site_count")),
                                 dp(d_out.get("pass")), dp(d_out.get("mit_sum")),
                                 dp(d_out.get("cfd_sum")))
        p = self.params(method, threshold, form, precision, flags, chunk)
        check(lib().issl_score_device(self._h, d_guides.ptr, n, byref(p), byref(res),
                                      getattr(stream, "handle", stream)))

    def bind_stream(self, stream) -> None:
        """Give `stream` (a hiprt.Stream or a raw hipStream_t) a workspace of
        its own on this index: score_device calls on it then overlap with
        calls on other streams (issl_index_bind_stream)."""
        check(lib().issl_index_bind_stream(self._h, getattr(stream, "handle", stream)))

    def unbind_stream(self, stream) -> None:
        """Wait for the stream's calls on this index and free its workspace."""
        check(lib().issl_index_unbind_stream(self._h, getattr(stream, "handle", stream)))

    def set_tuning(self, chunk: int = 1024, depth: int = 8, grid_per_cu: int = 6,
                   direct_depth: int = 8, heavy_kpost: int = 0, ord_depth: Optional[int] = None,
                   ord_grid: Optional[int] = None, ord_split: int = 0,
                   ovl_grid: int = 0) -> None:
        """heavy_kpost: guides above heavy_kpost*1024 postings use the tile kernel,
        the rest one wave per guide (0: every guide uses tiles; HEAVY_NEVER: none).
        ord_depth / ord_grid: the ORDERED form's shape (default: depth, grid_per_cu);
        ord_split: ORDERED guides above this many postings are split over tiles
        (0: the default threshold; SPLIT_NEVER: never); ovl_grid: tile-kernel
        workgroups per CU of launches on bound streams (0: the default)."""
        t = _capi.issl_tuning(chunk, 256, depth, grid_per_cu, direct_depth, heavy_kpost, 0.0,
                              depth if ord_depth is None else ord_depth,
                              grid_per_cu if ord_grid is None else ord_grid, ord_split, ovl_grid)
        check(lib().issl_set_tuning(self._h, byref(t)))

    def set_overlap_grid(self, ovl_grid: int) -> None:
        """Tile-kernel workgroups per CU for launches on streams bound to the
        index (0: the default); the rest of the tuning is kept."""
        t = self.tuning()
        t.ovl_grid = int(ovl_grid)
        check(lib().issl_set_tuning(self._h, byref(t)))


# ------------------------------------------------------------------ serving

class Request:
    """A submitted request: its ticket and the host buffers the results land in."""

    def __init__(self, server: "Server", ticket: int, results: Results):
        self.server = server
        self.ticket = ticket
        self.results = results
        self.status: Optional[int] = None

    def wait(self, timeout: float = -1.0) -> Results:
        """Block until done; raise IsslError if the request failed."""
        st = ctypes.c_int32()
        check(lib().issl_server_wait(self.server._h, self.ticket, timeout, byref(st)))
        self.server._inflight.pop(self.ticket, None)   # the server writes nothing more
        self.status = int(st.value)
        if self.status != _capi.ISSL_OK:
            raise _capi.IsslError(self.status, lib().issl_last_error().decode())
        return self.results

    def poll(self) -> bool:
        if self.status is not None:
            return True
        done, st = ctypes.c_int32(), ctypes.c_int32()
        check(lib().issl_server_poll(self.server._h, self.ticket, byref(done), byref(st)))
        if done.value:
            self.server._inflight.pop(self.ticket, None)
            self.status = int(st.value)
        return bool(done.value)


class Extractor:
    """Off-target site extraction on one GPU (issl_extract.h): the
    reference's ExtractOfftargets (ExtractOfftargets.cpp:19-356)."""

    def __init__(self, device: int = 0, chunk_bytes: int = 0):
        h = c_void_p()
        check(lib().issl_extractor_create(device, chunk_bytes, byref(h)))
        self._h = h
        self._distinct = 0

    def close(self) -> None:
        if self._h is not None and self._h.value:
            check(lib().issl_extractor_destroy(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_sequence(self, seq) -> None:
        """Scan one sequence: str, bytes or any byte buffer (not copied here)."""
        if isinstance(seq, str):
            seq = seq.encode()
        a = np.frombuffer(seq, dtype=np.uint8)
        check(lib().issl_extractor_add_sequence(self._h, _ptr(a), a.shape[0]))

    def add_path(self, path: str) -> None:
        check(lib().issl_extractor_add_path(self._h, os.fsencode(path)))

    def finish(self):
        """(sites, distinct sites) after sorting everything added so far."""
        a, b = c_uint64(), c_uint64()
        check(lib().issl_extractor_finish(self._h, byref(a), byref(b)))
        self._distinct = b.value
        return a.value, b.value

    def sites(self):
        """Distinct sites in text order: (signatures uint64, occurrences uint32)."""
        n = self._distinct
        sigs = np.zeros(max(n, 1), dtype=np.uint64)[:n]
        occ = np.zeros(max(n, 1), dtype=np.uint32)[:n]
        check(lib().issl_extractor_sites(self._h, _ptr(sigs), _ptr(occ)))
        return sigs, occ

    def write(self, path: str) -> None:
        check(lib().issl_extractor_write(self._h, os.fsencode(path)))

    def build_index(self, masks) -> "Index":
        masks = np.ascontiguousarray(masks, dtype=np.uint64)
        h = c_void_p()
        check(lib().issl_extractor_build_index(self._h, _ptr(masks), masks.shape[0], byref(h)))
        return Index(h.value)

    def stats(self):
        """(bases scanned, sites found) so far."""
        a, b = c_uint64(), c_uint64()
        check(lib().issl_extractor_stats(self._h, byref(a), byref(b)))
        return a.value, b.value


class Server:
    """Request loop over resident indexes on one device (issl_server_*)."""

    def __init__(self, device: int = 0, max_batch: int = 65536, hbm_reserve: int = 0,
                 slots: int = 3):
        cfg = _capi.issl_server_config(device, slots, max_batch, hbm_reserve)
        h = c_void_p()
        check(lib().issl_server_create(byref(cfg), byref(h)))
        self._h = h
        # Result buffers of requests not yet released (ticket -> Results): the
        # server writes into them until the request completes, whether or not
        # the caller still holds its Request.
        self._inflight = {}

    def close(self) -> None:
        if self._h is not None and self._h.value:
            check(lib().issl_server_destroy(self._h))
        self._h = None
        self._inflight.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_index(self, source) -> int:
        """Admit an index from a .issl path or an IndexArrays; returns its id."""
        iid = ctypes.c_int32()
        if isinstance(source, str):
            check(lib().issl_server_add_index(self._h, source.encode(), byref(iid)))
        else:
            ix = source
            check(lib().issl_server_add_index_arrays(
                self._h, ix.n_sites, ix.seq_len, ix.n_slices, _ptr(ix.sigs), _ptr(ix.masks),
                _ptr(ix.list_sizes), _ptr(ix.entries), byref(iid)))
        return int(iid.value)

    def remove_index(self, index_id: int) -> None:
        check(lib().issl_server_remove_index(self._h, index_id))

    def submit(self, index_id: int, guides, method: int = METHOD_MIT_AND_CFD,
               threshold: float = 0.0, form: int = FORM_BATCHED, precision: int = PREC_FP32C,
               flags: int = 0, chunk: int = 0, max_dist: int = 4) -> Request:
        g = np.ascontiguousarray(guides, dtype=np.uint64)
        n = g.shape[0]
        r = Results(np.zeros(n), np.zeros(n), np.zeros(n, np.uint64), np.zeros(n, np.uint64),
                    np.zeros(n, np.uint8), np.zeros(n), np.zeros(n))
        res = _capi.issl_results(_ptr(r.mit), _ptr(r.cfd), _ptr(r.occ_count),
                                 _ptr(r.site_count), _ptr(r.passed), _ptr(r.mit_sum),
                                 _ptr(r.cfd_sum))
        p = Index.params(method, threshold, form, precision, flags, chunk, max_dist)
        t = c_uint64()
        check(lib().issl_server_submit(self._h, index_id, _ptr(g), n, byref(p), byref(res),
                                       byref(t)))
        self._inflight[int(t.value)] = r
        return Request(self, int(t.value), r)

    def stats(self, reset: bool = False) -> _capi.issl_server_stats:
        out = _capi.issl_server_stats()
        check(lib().issl_server_get_stats(self._h, 1 if reset else 0, byref(out)))
        return out


def device_memory(device: int = 0):
    """(free, total) HBM bytes of a device."""
    fr, tot = c_uint64(), c_uint64()
    check(lib().issl_device_memory(device, byref(fr), byref(tot)))
    return int(fr.value), int(tot.value)
