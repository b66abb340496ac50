"""Calls on one index from several HIP streams (MI355X).

Every call on an index uses the index's one scoring workspace.  A call on
another stream than the previous call's first waits for that call's work:
calls issued back to back, with no host synchronization, on the index's own
stream and on callers' streams (one of them destroyed in between) must each
give the result they give alone.
"""
import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
from cracklingplusplus_gpu_amd import hiprt
from conftest import golden_index, load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def outputs(n):
    return {k: hiprt.DeviceArray(n, dt) for k, dt in
            (("mit", np.float64), ("cfd", np.float64), ("occ_count", np.uint64),
             ("site_count", np.uint64), ("pass", np.uint8))}


@pytest.mark.parametrize("precision", [issl.PREC_FP32C, issl.PREC_FP64])
@pytest.mark.parametrize("threshold", [0.0, 75.0])
def test_streams_share_the_workspace(gpu_available, precision, threshold):
    print(f"[gpu-workspace] precision {precision} threshold {threshold}", flush=True)
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    n = len(g)
    parts = [g, g[: n // 3], g[n // 2:], g[::-1].copy()]
    kw = dict(threshold=threshold, form=issl.FORM_BATCHED, precision=precision)
    s1, s2, s3 = hiprt.Stream(), hiprt.Stream(), hiprt.Stream()
    try:
        with issl.Index.from_arrays(hx) as ix:
            want = [ix.score(p, **kw) for p in parts]
            dgs = [hiprt.DeviceArray.from_host(p) for p in parts]
            # (part, stream) in call order; None is the index's own stream
            plan = [(0, s1), (1, None), (2, s2), (3, None), (0, None), (1, s1), (2, s3),
                    (3, s2), (0, s1), (1, None)]
            outs = []
            for i, (k, s) in enumerate(plan):
                o = outputs(len(parts[k]))
                for a in o.values():
                    a.fill_bytes(0xFF)
                outs.append(o)
            hiprt.synchronize()
            for i, (k, s) in enumerate(plan):
                ix.score_device(dgs[k], len(parts[k]), outs[i],
                                stream=None if s is None else s.handle, **kw)
                if s is s3:
                    s3.destroy()      # the next call waits for work on a destroyed stream
            hiprt.synchronize()
            s1.synchronize()
            s2.synchronize()
            for i, (k, s) in enumerate(plan):
                got = {key: a.to_host() for key, a in outs[i].items()}
                w = want[k]
                assert (bits(got["mit"]) == bits(w.mit)).all(), i
                assert (bits(got["cfd"]) == bits(w.cfd)).all(), i
                assert (got["occ_count"] == w.occ_count).all(), i
                assert (got["site_count"] == w.site_count).all(), i
                assert (got["pass"] == w.passed).all(), i
    finally:
        for s in (s1, s2, s3):
            s.destroy()


@pytest.mark.parametrize("precision", [issl.PREC_FP32C, issl.PREC_FP64])
def test_workspace_reuse_across_batch_sizes(gpu_available, precision):
    """One index scores batches that shrink and grow again: the scoring
    workspace is reused, and each result equals the same guides scored on a
    fresh index (the tile queue's counters stay valid for every size)."""
    print(f"[gpu-workspace] batch sizes, precision {precision}", flush=True)
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    kw = dict(threshold=0.0, form=issl.FORM_BATCHED, precision=precision)
    sizes = [len(g), 37, 1, len(g) // 2, 5, len(g)]
    with issl.Index.from_arrays(hx) as ix:
        got = [ix.score(g[:m], **kw) for m in sizes]
    for m, r in zip(sizes, got):
        with issl.Index.from_arrays(hx) as fresh:
            w = fresh.score(g[:m], **kw)
        assert (bits(r.mit) == bits(w.mit)).all(), m
        assert (bits(r.cfd) == bits(w.cfd)).all(), m
        assert (r.site_count == w.site_count).all(), m
