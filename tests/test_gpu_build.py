"""Device index construction (issl_index_build) against the host list builder
and the reference's golden vectors (MI355X).

The host builder restates ISSLCreateIndex.cpp:450-484 and is itself checked
against the reference's own index file; the device build must give the same
list sizes and posting entries bit for bit, and the same scores.
"""
import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
from conftest import GOLDEN_CASES, golden_index, load_golden

pytestmark = pytest.mark.gpu


def site_occ(hx):
    """Occurrence count per id from the first slice's postings."""
    e = hx.entries[: hx.n_sites]
    occ = np.zeros(hx.n_sites, dtype=np.uint32)
    occ[(e & np.uint64(0xFFFFFFFF)).astype(np.int64)] = (e >> np.uint64(32)).astype(np.uint32)
    return occ


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_device_build_equals_host_lists(gpu_available, name):
    hx = golden_index(name)
    occ = site_occ(hx)
    with issl.Index.build(hx.sigs, occ, hx.masks, seq_len=hx.seq_len) as ix:
        sizes, entries = ix.export()
        assert np.array_equal(sizes, hx.list_sizes)
        assert np.array_equal(entries, hx.entries)
        assert ix.validate() == 0
        g = load_golden(name)
        for m in (issl.METHOD_MIT_AND_CFD, issl.METHOD_CFD):
            r = ix.score(g["guides"], method=m, threshold=0.0)
            want = g[f"m{m}_t0_cfd"]
            assert (np.ascontiguousarray(r.cfd).view(np.uint64) == want.view(np.uint64)).all()


def test_device_build_of_generated_sites(gpu_available):
    sigs, occ, masks, guides, kinds = issl.synth(dict(
        n_sites=300_000, gc=0.41, repeat_frac=0.05, n_repeat_seeds=200, repeat_sub_p=0.1,
        occ_p_unique=0.8, occ_p_repeat=0.5, n_slices=8, slice_width=8, contiguous_masks=0,
        n_guides_site=100, n_guides_planted=50, n_guides_repeat=100, n_guides_random=50,
        seed=0xB0B))
    hx = issl.build_lists(sigs, occ, masks)
    with issl.Index.build(sigs, occ, masks) as ix, issl.Index.from_arrays(hx) as ref:
        sizes, entries = ix.export()
        assert np.array_equal(sizes, hx.list_sizes)
        assert np.array_equal(entries, hx.entries)
        a = ix.score(guides, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
        b = ref.score(guides, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
        assert (a.cfd.view(np.uint64) == b.cfd.view(np.uint64)).all()
        assert (a.site_count == b.site_count).all()
