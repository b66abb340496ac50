"""Streamed index (postings in host memory, slices streamed through HBM
buffers) against the resident index, the reference's golden vectors and the
CPU oracle (MI355X).

Tolerances:
  * PLAIN: bit-identical to the resident index and the golden vectors for
    every slice-buffer count (the guide state carried between windows keeps
    the reference's order of additions and exit point).
  * BATCHED FP64: within 1e-12 relative of the goldens, counts exact with
    the threshold off; FP32C within 1e-6.
"""
import itertools
import os

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import GOLDEN, golden_index, load_golden

pytestmark = pytest.mark.gpu

METHODS = range(5)
THRESHOLDS = (0, 75)
SLOT_BYTES_PER_SITE = 16


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def slot_bytes(hx, slots):
    return slots * SLOT_BYTES_PER_SITE * hx.n_sites


@pytest.fixture(scope="module")
def data(gpu_available):
    out = {}
    for name in ("legacy39_5x4", "tiny4k_5x4", "mid50k_10x10", "small200k_8x8"):
        print(f"[gpu-stream] loading {name}", flush=True)
        hx = golden_index(name)
        out[name] = (hx, load_golden(name))
    return out


@pytest.fixture(scope="module")
def resident(data):
    ixs = {name: issl.Index.from_arrays(hx) for name, (hx, _) in data.items()}
    yield ixs
    for ix in ixs.values():
        ix.close()


@pytest.mark.parametrize("slots", [1, 2, 3, 64])
@pytest.mark.parametrize("name", ["legacy39_5x4", "tiny4k_5x4", "mid50k_10x10", "small200k_8x8"])
def test_streamed_plain_bit_identical(data, resident, name, slots):
    hx, gd = data[name]
    g = gd["guides"]
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, slots)) as sx:
        n_slots, _, _ = sx.stream_stats()
        assert n_slots == min(slots, hx.n_slices)
        for m, t in itertools.product(METHODS, THRESHOLDS):
            print(f"[gpu-stream] plain {name} slots{slots} m{m} t{t}", flush=True)
            r = sx.score(g, method=m, threshold=float(t), form=issl.FORM_PLAIN)
            key = f"m{m}_t{t}"
            assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), key
            assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), key
            assert (r.passed == gd[key + "_pass"]).all(), key
            w = resident[name].score(g, method=m, threshold=float(t), form=issl.FORM_PLAIN)
            assert (bits(r.mit_sum) == bits(w.mit_sum)).all(), key
            assert (bits(r.cfd_sum) == bits(w.cfd_sum)).all(), key
            assert (r.occ_count == w.occ_count).all(), key
            assert (r.site_count == w.site_count).all(), key


@pytest.mark.parametrize("precision,rtol", [(issl.PREC_FP64, 1e-12), (issl.PREC_FP32C, 1e-6)])
@pytest.mark.parametrize("slots", [1, 3, 64])
@pytest.mark.parametrize("name", ["tiny4k_5x4", "mid50k_10x10", "small200k_8x8"])
def test_streamed_batched_vs_reference(data, name, slots, precision, rtol):
    hx, gd = data[name]
    g = gd["guides"]
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, slots)) as sx:
        for m in METHODS:
            print(f"[gpu-stream] batched {name} slots{slots} p{precision} m{m}", flush=True)
            r = sx.score(g, method=m, threshold=0.0, form=issl.FORM_BATCHED, precision=precision)
            key = f"m{m}_t0"
            np.testing.assert_allclose(r.mit, gd[key + "_mit"], rtol=rtol, atol=0)
            np.testing.assert_allclose(r.cfd, gd[key + "_cfd"], rtol=rtol, atol=0)
            o = orc.score(hx, g, m, 0.0)
            assert (r.site_count == o["site_count"]).all()
            assert (r.occ_count == o["occ_count"]).all()


@pytest.mark.parametrize("slots", [1, 2, 64])
@pytest.mark.parametrize("name", ["tiny4k_5x4", "mid50k_10x10", "small200k_8x8"])
def test_streamed_batched_exit_point(data, name, slots):
    """Threshold on: the exit is found in the right window and at the right site."""
    hx, gd = data[name]
    g = gd["guides"]
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, slots)) as sx:
        for m in METHODS:
            o = orc.score(hx, g, m, 75.0)
            r = sx.score(g, method=m, threshold=75.0, form=issl.FORM_BATCHED,
                         precision=issl.PREC_FP64)
            same = r.site_count == o["site_count"]
            assert same.mean() > 0.99
            np.testing.assert_allclose(r.mit_sum[same], o["mit_sum"][same], rtol=1e-12)
            np.testing.assert_allclose(r.cfd_sum[same], o["cfd_sum"][same], rtol=1e-12)


@pytest.mark.parametrize("form", [issl.FORM_PLAIN, issl.FORM_BATCHED])
@pytest.mark.parametrize("t", [0, 75])
def test_streamed_distinct_sites(data, form, t):
    hx, gd = data["small200k_8x8"]
    g = gd["guides"]
    o = orc.score(hx, g, 2, float(t), distinct=True)
    want = sum(bin(int(w)).count("1") for w in o["distinct"])
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, 2)) as sx:
        sx.distinct_reset()
        sx.score(g, method=2, threshold=float(t), form=form, flags=issl.FLAG_DISTINCT)
        assert sx.distinct_count() == want


def test_streamed_upload_accounting(data):
    """Fewer buffers than slices: every call streams the whole index once;
    buffers for every slice: the slices are copied once and stay."""
    hx, gd = data["mid50k_10x10"]
    g = gd["guides"]
    per_pass = hx.n_slices * hx.n_sites * 8
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, 4)) as sx:
        for call in range(3):
            sx.score(g, form=issl.FORM_PLAIN)
        slots, up, win = sx.stream_stats()
        assert slots == 4
        assert up == 3 * per_pass
        assert win == 3 * 5  # windows of 2 slices
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, hx.n_slices)) as sx:
        a = sx.score(g, form=issl.FORM_PLAIN)
        b = sx.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
        c = sx.score(g, form=issl.FORM_PLAIN)
        slots, up, win = sx.stream_stats()
        assert slots == hx.n_slices
        assert up == per_pass
        assert win == 3
        assert (bits(a.mit) == bits(c.mit)).all()
        np.testing.assert_allclose(b.mit, a.mit, rtol=1e-12)


def test_streamed_split_batches(data, resident):
    """More guides than one batched launch takes: the carried state is offset
    per launch."""
    hx, gd = data["mid50k_10x10"]
    g = np.tile(gd["guides"], (300_000 + len(gd["guides"]) - 1) // len(gd["guides"]))[:300_000]
    want = resident["mid50k_10x10"].score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, 3)) as sx:
        r = sx.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
        np.testing.assert_allclose(r.mit, want.mit, rtol=1e-12)
        np.testing.assert_allclose(r.cfd, want.cfd, rtol=1e-12)
        assert (r.site_count == want.site_count).all()
        p = sx.score(g[:20_000], form=issl.FORM_PLAIN)
        w = resident["mid50k_10x10"].score(g[:20_000], form=issl.FORM_PLAIN)
        assert (bits(p.mit) == bits(w.mit)).all()


def test_open_streamed_file(data, tmp_path):
    hx, gd = data["tiny4k_5x4"]
    path = os.path.join(GOLDEN, "tiny4k_5x4.issl")
    g = gd["guides"]
    with issl.Index.open_streamed(path, slot_bytes=slot_bytes(hx, 2)) as sx:
        for m, t in itertools.product(METHODS, THRESHOLDS):
            r = sx.score(g, method=m, threshold=float(t), form=issl.FORM_PLAIN)
            key = f"m{m}_t{t}"
            assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), key
            assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), key
    with issl.Index.open_streamed(path) as sx:   # budget from the free memory: all resident
        assert sx.stream_stats()[0] == hx.n_slices
        r = sx.score(g, method=2, threshold=75.0, form=issl.FORM_PLAIN)
        assert (r.passed == gd["m2_t75_pass"]).all()


def test_streamed_rejects_bad_ids_before_scoring(data, tmp_path):
    """An entry whose id is >= N is found while it is copied: the call fails
    with a format error and no kernel reads it."""
    hx, gd = data["tiny4k_5x4"]
    bad = issl.IndexArrays(hx.sigs.copy(), hx.masks.copy(), hx.list_sizes.copy(),
                           hx.entries.copy(), hx.seq_len)
    last = (hx.n_slices - 1) * hx.n_sites + 7
    bad.entries[last] = (bad.entries[last] & np.uint64(0xFFFFFFFF00000000)) | np.uint64(hx.n_sites + 5)
    with issl.Index.from_arrays_streamed(bad, slot_bytes=slot_bytes(hx, 2)) as sx:
        with pytest.raises(issl.IsslError, match="ids >= N"):
            sx.score(gd["guides"], form=issl.FORM_PLAIN)


def test_streamed_resident_only_operations(data):
    hx, gd = data["tiny4k_5x4"]
    with issl.Index.from_arrays_streamed(hx, slot_bytes=slot_bytes(hx, 2)) as sx:
        with pytest.raises(issl.IsslError, match="resident"):
            sx.validate()
        with pytest.raises(issl.IsslError, match="resident"):
            sx.export()
        with pytest.raises(issl.IsslError, match="resident"):
            sx.candidates(gd["guides"][:4])
        with pytest.raises(issl.IsslError, match="resident"):
            sx.autotune(gd["guides"][:64])
    with pytest.raises(issl.IsslError, match="no slice buffer"):
        issl.Index.from_arrays_streamed(hx, slot_bytes=SLOT_BYTES_PER_SITE * hx.n_sites - 1)
