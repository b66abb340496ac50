"""List-residual site words against the signature-inline words (MI355X).

The BATCHED tile kernel streams, per posting, either the signature-inline word
(8 B: signature | occurrences << 40) or the list-residual word (4 B: the codes
at the positions outside the slice, then a clipped occurrence count).  Both
feed the same candidates, in the same order, to the same terms, so every
output must be the same bits.  Occurrence counts at or above the residual
word's escape (2^bits - 1) are read from the posting; the escape cases below
plant counts on both sides of it, and of the inline word's 2^24 - 1.
"""
import itertools

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import golden_index, load_golden
from cracklingplusplus_gpu_amd import workloads

pytestmark = pytest.mark.gpu

CASES = ("tiny4k_5x4", "small200k_8x8", "mid50k_10x10")
# (built, occurrence bits) by the narrowest slice: 5x4 has 4 positions (too
# narrow), 8x8 leaves 12 residual positions (8 occurrence bits), 10x10 leaves
# 10 (12 bits).
EXPECT = {"tiny4k_5x4": (False, 0), "small200k_8x8": (True, 8), "mid50k_10x10": (True, 12)}


@pytest.fixture(scope="module")
def cases(gpu_available):
    out = {}
    for name in CASES:
        print(f"[gpu-residual] opening {name}", flush=True)
        hx = golden_index(name)
        out[name] = (hx, issl.Index.from_arrays(hx), load_golden(name))
    yield out
    for _, ix, _ in out.values():
        ix.close()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_bits(a, b, what):
    assert (bits(a.mit) == bits(b.mit)).all(), what
    assert (bits(a.cfd) == bits(b.cfd)).all(), what
    assert (bits(a.mit_sum) == bits(b.mit_sum)).all(), what
    assert (bits(a.cfd_sum) == bits(b.cfd_sum)).all(), what
    assert (a.occ_count == b.occ_count).all(), what
    assert (a.site_count == b.site_count).all(), what
    assert (a.passed == b.passed).all(), what


def both_layouts(ix, g, **kw):
    ix.set_residual(True)
    try:
        res = ix.score(g, form=issl.FORM_BATCHED, **kw)
        ix.set_residual(False)
        inl = ix.score(g, form=issl.FORM_BATCHED, **kw)
    finally:
        ix.set_residual(False)
    return res, inl


@pytest.mark.parametrize("name", CASES)
def test_residual_built_where_it_fits(cases, name):
    _, ix, _ = cases[name]
    built, used, occ_bits = ix.residual()
    assert (built, occ_bits) == EXPECT[name]
    assert used is False          # a new index streams the signature-inline words
    ix.set_residual(True)
    assert ix.residual()[1] == built
    ix.set_residual(False)
    assert ix.residual()[1] is False


@pytest.mark.parametrize("precision", [issl.PREC_FP64, issl.PREC_FP32C])
@pytest.mark.parametrize("name", ["small200k_8x8", "mid50k_10x10"])
def test_residual_same_bits_as_inline(cases, name, precision):
    hx, ix, gd = cases[name]
    g = gd["guides"]
    for m, t in itertools.product(range(5), (0, 75)):
        print(f"[gpu-residual] {name} p{precision} m{m} t{t}", flush=True)
        res, inl = both_layouts(ix, g, method=m, threshold=float(t), precision=precision)
        same_bits(res, inl, f"{name} m{m} t{t}")
        key = f"m{m}_t{t}"
        if t == 0:
            rtol = 1e-12 if precision == issl.PREC_FP64 else 1e-6
            np.testing.assert_allclose(res.cfd, gd[key + "_cfd"], rtol=rtol)
            np.testing.assert_allclose(res.mit, gd[key + "_mit"], rtol=rtol)


@pytest.mark.parametrize("name", ["small200k_8x8", "mid50k_10x10"])
def test_residual_tile_shapes(cases, name):
    """Short and long tiles, shallow and deep rounds: the same bits as inline."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    try:
        for chunk, depth, gpc in [(64, 16, 1), (512, 4, 8), (4096, 8, 4)]:
            ix.set_tuning(chunk=chunk, depth=depth, grid_per_cu=gpc, heavy_kpost=0)
            res, inl = both_layouts(ix, g, method=issl.METHOD_MIT_AND_CFD,
                                    precision=issl.PREC_FP32C)
            same_bits(res, inl, f"{name} chunk {chunk} depth {depth}")
    finally:
        ix.set_tuning()


# Counts straddling the 8-bit and 12-bit escapes (255, 4095) and the inline
# word's 2^24 - 1.
SPECIAL_OCC = np.array([1, 2, 3, 254, 255, 256, 1000, 4094, 4095, 4096, 70_000,
                        (1 << 24) - 2, (1 << 24) - 1, 1 << 24, (1 << 24) + 3, 3_000_000_000],
                       dtype=np.uint32)


@pytest.mark.parametrize("name", ["small200k_8x8", "mid50k_10x10"])
def test_residual_occurrence_escape(gpu_available, name):
    sigs, occ, masks, guides, _ = workloads.make(name, n_sites=30_000)
    rng = np.random.default_rng(11)
    occ = SPECIAL_OCC[rng.integers(0, len(SPECIAL_OCC), size=len(occ))].astype(occ.dtype)
    hx = issl.build_lists(sigs, occ, masks)
    g = guides[:400]
    with issl.Index.from_arrays(hx) as ix:
        assert ix.residual()[0]
        for m, t in [(issl.METHOD_MIT_AND_CFD, 0.0), (issl.METHOD_CFD, 0.0), (issl.METHOD_MIT, 75.0)]:
            o = orc.score(hx, g, m, t)
            for prec, rtol in [(issl.PREC_FP64, 1e-12), (issl.PREC_FP32C, 1e-6)]:
                res, inl = both_layouts(ix, g, method=m, threshold=t, precision=prec)
                same_bits(res, inl, f"{name} m{m} t{t} p{prec}")
                if t == 0.0:
                    assert (res.occ_count == o["occ_count"]).all()
                    assert (res.site_count == o["site_count"]).all()
                    np.testing.assert_allclose(res.mit_sum, o["mit_sum"], rtol=rtol)
                    np.testing.assert_allclose(res.cfd_sum, o["cfd_sum"], rtol=rtol)
        # distinct sites read each hit's id from its posting
        od = orc.score(hx, g, issl.METHOD_MIT_AND_CFD, 0.0, distinct=True)
        want = sum(bin(int(w)).count("1") for w in od["distinct"])
        for use in (True, False):
            ix.set_residual(use)
            ix.distinct_reset()
            ix.score(g, form=issl.FORM_BATCHED, flags=issl.FLAG_DISTINCT)
            assert ix.distinct_count() == want
        ix.set_residual(False)
