"""hipGraph capture of the scoring launch sequence (ISSL_FLAG_GRAPH), MI355X.

The first flagged issl_score_device call runs the sequence and captures it;
later calls with the same buffers replay the graph.  Each replay is checked
against the eager result, with eager work of other shapes and host
synchronizations between the replays, and the captured sequence holds no
memset node (the kernels clear nothing through the runtime).
"""
import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
from cracklingplusplus_gpu_amd import hiprt
from conftest import golden_index, load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("form,precision", [(issl.FORM_BATCHED, issl.PREC_FP32C),
                                            (issl.FORM_BATCHED, issl.PREC_FP64),
                                            (issl.FORM_PLAIN, issl.PREC_FP64)])
@pytest.mark.parametrize("threshold", [0.0, 75.0])
def test_graph_replays_match_eager(gpu_available, form, precision, threshold):
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    n = len(g)
    with issl.Index.from_arrays(hx) as ix:
        want = ix.score(g, threshold=threshold, form=form, precision=precision)
        dg = hiprt.DeviceArray.from_host(g)
        out = {k: hiprt.DeviceArray(n, dt) for k, dt in
               (("mit", np.float64), ("cfd", np.float64), ("occ_count", np.uint64),
                ("site_count", np.uint64), ("pass", np.uint8))}
        kw = dict(threshold=threshold, form=form, precision=precision, flags=issl.FLAG_GRAPH)
        other = g[: n // 3]
        for rep in range(5):
            for a in out.values():                       # stale values must be overwritten
                a.fill_bytes(0xFF)
            hiprt.synchronize()      # the fills run on the null stream, the index on its own
            ix.score_device(dg, n, out, **kw)            # rep 0 runs eagerly and captures
            hiprt.synchronize()
            got = {k: a.to_host() for k, a in out.items()}
            assert (bits(got["mit"]) == bits(want.mit)).all(), rep
            assert (bits(got["cfd"]) == bits(want.cfd)).all(), rep
            assert (got["occ_count"] == want.occ_count).all(), rep
            assert (got["site_count"] == want.site_count).all(), rep
            assert (got["pass"] == want.passed).all(), rep
            # eager work of another shape on the same index between replays
            r = ix.score(other, threshold=threshold, form=form, precision=precision)
            assert (bits(r.cfd) == bits(want.cfd[: n // 3])).all(), rep


@pytest.mark.parametrize("threshold", [0.0, 75.0])
def test_graph_replays_between_one_tile_launches(gpu_available, threshold):
    """With a chunk that covers every guide (one tile each), replays
    interleaved with eager launches of another batch size must still match
    the eager result, and the eager launches each other."""
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    n = len(g)
    with issl.Index.from_arrays(hx) as ix:
        ix.set_tuning(chunk=1 << 20, depth=8, grid_per_cu=4, heavy_kpost=0)
        want = ix.score(g, threshold=threshold, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
        dg = hiprt.DeviceArray.from_host(g)
        out = {k: hiprt.DeviceArray(n, dt) for k, dt in
               (("mit", np.float64), ("cfd", np.float64), ("occ_count", np.uint64),
                ("site_count", np.uint64), ("pass", np.uint8))}
        kw = dict(threshold=threshold, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C,
                  flags=issl.FLAG_GRAPH)
        other = g[: n // 3]
        for rep in range(5):
            for a in out.values():
                a.fill_bytes(0xFF)
            hiprt.synchronize()
            ix.score_device(dg, n, out, **kw)
            hiprt.synchronize()
            got = {k: a.to_host() for k, a in out.items()}
            assert (bits(got["cfd"]) == bits(want.cfd)).all(), rep
            assert (bits(got["mit"]) == bits(want.mit)).all(), rep
            assert (got["site_count"] == want.site_count).all(), rep
            assert (got["pass"] == want.passed).all(), rep
            # one or two eager one-tile launches of another size
            for _ in range(1 + rep % 2):
                r = ix.score(other, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
                w0 = ix.score(other, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
                assert (bits(r.cfd) == bits(w0.cfd)).all(), rep
