"""The allocation-free entry point (issl_score_device_noalloc) through the C
ABI against issl_score_device: the same bytes for every form and precision,
on two calls in a row, the second with fewer guides (MI355X).  Calls beyond
its workspace are refused with ISSL_ERR_CAPACITY and leave it usable."""
from ctypes import byref

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
from cracklingplusplus_gpu_amd import _capi, hiprt
from conftest import golden_index, load_golden

pytestmark = pytest.mark.gpu

COLUMNS = (("mit", np.float64), ("cfd", np.float64), ("mit_sum", np.float64),
           ("cfd_sum", np.float64), ("occ_count", np.uint64), ("site_count", np.uint64),
           ("pass", np.uint8))
FORMS = {"batched_fp64": (issl.FORM_BATCHED, issl.PREC_FP64),
         "batched_fp32c": (issl.FORM_BATCHED, issl.PREC_FP32C),
         "ordered": (issl.FORM_ORDERED, issl.PREC_FP64),
         "plain": (issl.FORM_PLAIN, issl.PREC_FP64)}


def call(entry, ix, d_g, n, params):
    """One scoring call through the named C entry point into fresh device
    columns (filled with 0xA5 first, so unwritten entries show)."""
    out = {}
    for name, dt in COLUMNS:
        out[name] = hiprt.DeviceArray(max(n, 1), dt)
        out[name].fill_bytes(0xA5)
    res = _capi.issl_results(out["mit"].ptr, out["cfd"].ptr, out["occ_count"].ptr,
                             out["site_count"].ptr, out["pass"].ptr, out["mit_sum"].ptr,
                             out["cfd_sum"].ptr)
    rc = getattr(issl.lib(), entry)(ix.handle, d_g.ptr, n, byref(params), byref(res), None)
    issl.check(rc)
    hiprt.synchronize()
    return {k: v.to_host()[:n] for k, v in out.items()}


@pytest.fixture(scope="module")
def indexes(gpu_available):
    out = {}
    for name in ("mid50k_10x10", "uniform1m_5x4"):
        hx = golden_index(name)
        out[name] = (issl.Index.from_arrays(hx), load_golden(name)["guides"])
    yield out
    for ix, _ in out.values():
        ix.close()


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("name", ["mid50k_10x10", "uniform1m_5x4"])
def test_noalloc_matches_score_device(indexes, name, form):
    ix, g = indexes[name]
    f, prec = FORMS[form]
    d_g = hiprt.DeviceArray.from_host(np.ascontiguousarray(g))
    for t in (0.0, 75.0):
        p = issl.Index.params(method=issl.METHOD_MIT_AND_CFD, threshold=t, form=f, precision=prec)
        for n in (len(g), 137):            # two calls in a row, the second smaller
            print(f"[no-alloc] {name} {form} t{t} n{n}", flush=True)
            want = call("issl_score_device", ix, d_g, n, p)
            got = call("issl_score_device_noalloc", ix, d_g, n, p)
            for k, _ in COLUMNS:
                assert got[k].tobytes() == want[k].tobytes(), (name, form, t, n, k)


def test_noalloc_capacity_and_recovery(indexes):
    ix, g = indexes["mid50k_10x10"]
    big = np.resize(g, 200_000)
    d_big = hiprt.DeviceArray.from_host(big)
    p = issl.Index.params(form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
    with pytest.raises(issl.IsslError) as e:
        call("issl_score_device_noalloc", ix, d_big, big.shape[0], p)
    assert e.value.code == _capi.ISSL_ERR_CAPACITY and "at most" in str(e.value)
    # the limit it names is accepted, and the PLAIN form needs no workspace
    limit = int(str(e.value).split("at most ")[1].split()[0])
    assert 1000 < limit < 200_000
    n = min(limit, big.shape[0])
    want = call("issl_score_device", ix, d_big, n, p)
    got = call("issl_score_device_noalloc", ix, d_big, n, p)
    assert got["cfd"].tobytes() == want["cfd"].tobytes()
    pp = issl.Index.params(form=issl.FORM_PLAIN)
    got = call("issl_score_device_noalloc", ix, d_big, 20_000, pp)
    want = call("issl_score_device", ix, d_big, 20_000, pp)
    assert got["mit"].tobytes() == want["mit"].tobytes()
    with pytest.raises(issl.IsslError):
        call("issl_score_device_noalloc", ix, d_big, 10,
             issl.Index.params(form=issl.FORM_BATCHED, flags=issl.FLAG_GRAPH))
