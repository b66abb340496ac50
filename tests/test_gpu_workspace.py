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


def same_as(got, w, tag):
    assert (bits(got["mit"]) == bits(w.mit)).all(), tag
    assert (bits(got["cfd"]) == bits(w.cfd)).all(), tag
    assert (got["occ_count"] == w.occ_count).all(), tag
    assert (got["site_count"] == w.site_count).all(), tag
    assert (got["pass"] == w.passed).all(), tag


FORMS = [(issl.FORM_BATCHED, issl.PREC_FP64), (issl.FORM_BATCHED, issl.PREC_FP32C),
         (issl.FORM_ORDERED, issl.PREC_FP64), (issl.FORM_PLAIN, issl.PREC_FP64)]


@pytest.mark.parametrize("form,precision", FORMS)
@pytest.mark.parametrize("threshold", [0.0, 75.0])
def test_bound_streams_overlap_with_own_workspaces(gpu_available, form, precision, threshold):
    """Two bound streams and the index's own stream, calls issued back to back
    with no host synchronization (each bound stream scoring in its own
    workspace, so the calls run at once): every call gives the result it gives
    alone.  Then one stream is unbound and scores through the shared workspace
    again."""
    print(f"[gpu-workspace] bound streams form {form} precision {precision} "
          f"threshold {threshold}", flush=True)
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    n = len(g)
    parts = [g, g[: n // 3], g[n // 2:], g[::-1].copy()]
    kw = dict(threshold=threshold, form=form, precision=precision)
    s1, s2 = hiprt.Stream(), hiprt.Stream()
    try:
        with issl.Index.from_arrays(hx) as ix:
            want = [ix.score(p, **kw) for p in parts]
            dgs = [hiprt.DeviceArray.from_host(p) for p in parts]
            ix.bind_stream(s1)
            ix.bind_stream(s2)
            ix.bind_stream(s1)    # binding again does nothing
            ix.set_overlap_grid(2)   # bound streams' tile kernels at 2 workgroups per CU
            plan = [(k % 4, s) for k, s in enumerate([s1, s2, None] * 6)]
            outs = []
            for k, s in plan:
                o = outputs(len(parts[k]))
                for a in o.values():
                    a.fill_bytes(0xFF)
                outs.append(o)
            hiprt.synchronize()
            for i, (k, s) in enumerate(plan):
                ix.score_device(dgs[k], len(parts[k]), outs[i], stream=s, **kw)
            ix.unbind_stream(s1)   # waits for s1's calls
            hiprt.synchronize()
            s2.synchronize()
            for i, (k, s) in enumerate(plan):
                same_as({key: a.to_host() for key, a in outs[i].items()}, want[k], i)
            # s1 is unbound: it shares the index's workspace again
            o = outputs(n)
            ix.score_device(dgs[0], n, o, stream=s1, **kw)
            s1.synchronize()
            same_as({key: a.to_host() for key, a in o.items()}, want[0], "unbound")
            ix.unbind_stream(s2)
    finally:
        s1.destroy()
        s2.destroy()


def test_bind_stream_errors(gpu_available):
    print("[gpu-workspace] bind errors", flush=True)
    hx = golden_index("mid50k_10x10")
    streams = [hiprt.Stream() for _ in range(9)]
    try:
        with issl.Index.from_arrays(hx) as ix:
            with pytest.raises(issl.IsslError):
                ix.unbind_stream(streams[0])          # never bound
            for s in streams[:8]:
                ix.bind_stream(s)
            with pytest.raises(issl.IsslError) as e:
                ix.bind_stream(streams[8])            # at most 8 per index
            assert e.value.code == issl._capi.ISSL_ERR_CAPACITY
            for s in streams[:8]:
                ix.unbind_stream(s)
            ix.bind_stream(streams[8])
            # the bound streams' tile-kernel grid is kept beside the tuning
            ix.set_overlap_grid(2)
            assert ix.tuning().ovl_grid == 2
            with pytest.raises(issl.IsslError):
                ix.set_overlap_grid(33)
            assert ix.tuning().ovl_grid == 2
            # closing the index with a stream still bound is allowed
    finally:
        for s in streams:
            s.destroy()


def test_graph_replays_on_bound_streams(gpu_available):
    """ISSL_FLAG_GRAPH on two bound streams: each stream's captured sequence
    is kept beside the other's and every replay equals the eager result."""
    print("[gpu-workspace] graphs on bound streams", flush=True)
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    n = len(g)
    kw = dict(threshold=0.0, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
    s1, s2 = hiprt.Stream(), hiprt.Stream()
    try:
        with issl.Index.from_arrays(hx) as ix:
            want = ix.score(g, **kw)
            dg = hiprt.DeviceArray.from_host(g)
            ix.bind_stream(s1)
            ix.bind_stream(s2)
            o1, o2 = outputs(n), outputs(n)
            for r in range(5):
                for o in (o1, o2):
                    for a in o.values():
                        a.fill_bytes(0xFF)
                hiprt.synchronize()
                ix.score_device(dg, n, o1, stream=s1, flags=issl.FLAG_GRAPH, **kw)
                ix.score_device(dg, n, o2, stream=s2, flags=issl.FLAG_GRAPH, **kw)
                s1.synchronize()
                s2.synchronize()
                same_as({key: a.to_host() for key, a in o1.items()}, want, ("s1", r))
                same_as({key: a.to_host() for key, a in o2.items()}, want, ("s2", r))
                ix.score(g[: n // 2], **kw)   # eager work between replays
            ix.unbind_stream(s1)
            ix.unbind_stream(s2)
    finally:
        s1.destroy()
        s2.destroy()
