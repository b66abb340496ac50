"""Serving loop (issl_server_*): isolation under concurrent load, continuous
batching, admission by free memory, failure containment (MI355X).

Results of a served request must be bit-identical to scoring the same guides
alone through issl_score with the same parameters: each guide's result depends
on the guide only, so merging requests must be invisible.
"""
import threading

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
from conftest import golden_index, load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def served(gpu_available):
    names = ["tiny4k_5x4", "mid50k_10x10"]
    hx = {n: golden_index(n) for n in names}
    gd = {n: load_golden(n) for n in names}
    srv = issl.Server(max_batch=4096)
    ids = {n: srv.add_index(hx[n]) for n in names}
    direct = {n: issl.Index.from_arrays(hx[n]) for n in names}
    yield srv, ids, direct, gd
    srv.close()
    for ix in direct.values():
        ix.close()


def same(a, b):
    return ((bits(a.mit) == bits(b.mit)).all() and (bits(a.cfd) == bits(b.cfd)).all()
            and (a.occ_count == b.occ_count).all() and (a.site_count == b.site_count).all()
            and (a.passed == b.passed).all())


@pytest.mark.parametrize("precision", [issl.PREC_FP32C, issl.PREC_FP64])
def test_concurrent_requests_match_direct_scoring(served, precision):
    srv, ids, direct, gd = served
    rng = np.random.default_rng(11)
    jobs = []
    for k in range(48):
        name = ["tiny4k_5x4", "mid50k_10x10"][k % 2]
        g = gd[name]["guides"]
        lo = int(rng.integers(0, len(g)))
        n = int(rng.integers(1, 300))
        t = [0.0, 75.0][(k // 2) % 2]
        jobs.append((name, np.resize(g[lo:], n) if lo + n > len(g) else g[lo:lo + n], t))
    got = [None] * len(jobs)

    def client(idx):
        reqs = []
        for j in idx:
            name, g, t = jobs[j]
            reqs.append((j, srv.submit(ids[name], g, threshold=t, precision=precision)))
        for j, r in reqs:
            got[j] = r.wait(timeout=60.0)

    threads = [threading.Thread(target=client, args=(list(range(i, len(jobs), 4)),))
               for i in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for j, (name, g, t) in enumerate(jobs):
        want = direct[name].score(g, threshold=t, form=issl.FORM_BATCHED, precision=precision)
        assert got[j] is not None and same(got[j], want), j


def test_queued_requests_are_merged(gpu_available):
    """Requests queued behind a busy slot share the next launch."""
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    with issl.Index.from_arrays(hx) as ix, issl.Server(slots=1, max_batch=1 << 16) as srv:
        iid = srv.add_index(hx)
        srv.submit(iid, g).wait(timeout=60.0)        # warm the slot
        srv.stats(reset=True)
        busy = srv.submit(iid, np.tile(g, 40))       # keeps the only slot busy
        reqs = [srv.submit(iid, g[i:i + 10]) for i in range(0, len(g), 10)]
        busy.wait(timeout=60.0)
        for r in reqs:
            r.wait(timeout=60.0)
        st = srv.stats(reset=True)
        assert st.requests_done == len(reqs) + 1 and st.requests_failed == 0
        assert st.launches < len(reqs)               # the small requests shared launches
        assert st.guides_done == 41 * len(g)
        assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
        assert st.resident_bytes > 0 and st.hbm_total >= st.hbm_free
        want = ix.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
        for k, r in enumerate(reqs):
            lo = 10 * k
            assert (bits(r.results.cfd) == bits(want.cfd[lo:lo + 10])).all(), k


def test_bad_request_fails_alone(served):
    srv, ids, direct, gd = served
    g = gd["tiny4k_5x4"]["guides"][:50]
    good = srv.submit(ids["tiny4k_5x4"], g)
    with pytest.raises(issl.IsslError):
        srv.submit(ids["tiny4k_5x4"], g, max_dist=3)
    with pytest.raises(issl.IsslError):
        srv.submit(12345, g)
    r = good.wait(timeout=60.0)
    assert same(r, direct["tiny4k_5x4"].score(g, form=issl.FORM_BATCHED,
                                              precision=issl.PREC_FP32C))


def test_admission_by_free_memory(served):
    srv, ids, direct, gd = served
    free, total = issl.device_memory(0)
    assert 0 < free <= total
    hx = golden_index("tiny4k_5x4")
    with issl.Server(hbm_reserve=total) as tight:
        with pytest.raises(issl.IsslError, match="admission"):
            tight.add_index(hx)
    with issl.Server(hbm_reserve=1 << 20) as roomy:
        iid = roomy.add_index(hx)
        r = roomy.submit(iid, gd["tiny4k_5x4"]["guides"][:20]).wait(timeout=60.0)
        assert r.mit.shape == (20,)


def test_remove_index_fails_only_its_requests(gpu_available):
    hx = golden_index("tiny4k_5x4")
    gd = load_golden("tiny4k_5x4")
    g = gd["guides"]
    with issl.Server() as srv:
        a = srv.add_index(hx)
        b = srv.add_index(hx)
        keep = [srv.submit(b, g[:100]) for _ in range(5)]
        srv.remove_index(a)
        with pytest.raises(issl.IsslError):
            srv.submit(a, g[:10])
        for r in keep:
            r.wait(timeout=60.0)


def test_large_requests_split_over_slots_match_direct(gpu_available):
    """A request larger than a slot runs as several launches; others interleave."""
    hx = golden_index("mid50k_10x10")
    g = load_golden("mid50k_10x10")["guides"]
    with issl.Index.from_arrays(hx) as ix, issl.Server(max_batch=96, slots=3) as srv:
        iid = srv.add_index(hx)
        big = srv.submit(iid, g)                       # 500 guides -> 6 launches
        small = [srv.submit(iid, g[i:i + 7], threshold=75.0) for i in range(0, 70, 7)]
        r = big.wait(timeout=60.0)
        assert same(r, ix.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C))
        for k, q in enumerate(small):
            lo = 7 * k
            want = ix.score(g[lo:lo + 7], threshold=75.0, form=issl.FORM_BATCHED,
                            precision=issl.PREC_FP32C)
            assert same(q.wait(timeout=60.0), want), k
        st = srv.stats()
        assert st.launches >= 6 and st.requests_failed == 0
        assert st.peak_inflight >= 2      # launches overlapped on the device


def test_failed_split_request_completes_after_its_launches(gpu_available):
    """A request split over many launches whose index is removed while some of
    them are in flight fails, but completes only once no launch holds any of
    its guides: after wait() its buffers are never written again, and the
    guides that did come back carry the direct scoring's results."""
    import time
    hx = golden_index("mid50k_10x10")
    g = np.resize(load_golden("mid50k_10x10")["guides"], 200_000)
    with issl.Index.from_arrays(hx) as ix, issl.Server(max_batch=256, slots=3) as srv:
        iid = srv.add_index(hx)
        req = srv.submit(iid, g, precision=issl.PREC_FP64)   # 782 launches
        res = req.results
        del req                                # the server keeps the buffers alive
        srv.remove_index(iid)                  # fails the rest; in-flight launches finish
        ticket = next(iter(srv._inflight))
        r = issl.Request(srv, ticket, res)
        with pytest.raises(issl.IsslError) as e:
            r.wait(timeout=60.0)
        assert e.value.code == issl._capi.ISSL_ERR_STATE
        assert ticket not in srv._inflight
        snap = {k: getattr(res, k).copy() for k in ("mit", "cfd", "occ_count", "passed")}
        time.sleep(0.2)
        for k, v in snap.items():
            assert np.array_equal(getattr(res, k), v), k
        back = res.mit != 0                    # a scored guide has 10000/(100+S) > 0
        assert int(back.sum()) < g.shape[0]
        if back.any():
            want = ix.score(g[back], form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
            assert (bits(res.mit[back]) == bits(want.mit)).all()
            assert (bits(res.cfd[back]) == bits(want.cfd)).all()
        st = srv.stats()
        assert st.requests_failed == 1
