"""Parity of the benchmark configurations beyond the 10,000-guide C1 step, at
full size, on an MI355X:

* 100,000 guides on the 100 M-site index in one launch (genome100m_10x10_q100k,
  30 % repeat neighbourhoods): ORDERED bit-exact against the oracle on a
  sample of every kind; BATCHED FP64 (1e-12) and FP32C (1e-6) against ORDERED
  for every guide with identical counts, under both site-word layouts forced
  (the list-residual words are what the autotuner picks at this size).
* 300,000,000 sites, 50,000 guides (genome300m_10x10): S * N = 3e9 postings,
  past 2^31, so the 64-bit posting offsets are in play.  PLAIN and ORDERED
  bit-exact against the oracle on a sample, threshold off and on; BATCHED
  FP64 / FP32C against ORDERED for every guide under both layouts; site
  counts against a brute force over all 300 M sites.
* Both indexes resident in one server: mixed-size requests against either
  index, each result bit-identical to scoring its guides directly.
* A heavy tail (genome100m_10x10_heavy): 1 % of 10,000 guides visit
  100-1000x the median posting volume; the exact forms against the oracle on
  a sample holding heavy guides, BATCHED against ORDERED for every guide.

Reference: lib/ISSLScoringModule.cpp:204-462.
"""
import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from cracklingplusplus_gpu_amd import workloads

pytestmark = pytest.mark.gpu

AND = issl.METHOD_MIT_AND_CFD


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def threads():
    import os
    return max(1, min(16, len(os.sched_getaffinity(0))))


def sample(kinds, per_kind, seed):
    rng = np.random.default_rng(seed)
    pick = []
    for k in np.unique(kinds):
        idx = np.nonzero(kinds == k)[0]
        pick.extend(rng.choice(idx, size=min(per_kind, idx.size), replace=False).tolist())
    return np.sort(np.array(pick))


def same_bits(a, b):
    return ((bits(a.mit) == bits(b.mit)).all() and (bits(a.cfd) == bits(b.cfd)).all() and
            (bits(a.mit_sum) == bits(b.mit_sum)).all() and (bits(a.cfd_sum) == bits(b.cfd_sum)).all()
            and (a.occ_count == b.occ_count).all() and (a.site_count == b.site_count).all() and
            (a.passed == b.passed).all())


def exact_vs_oracle(hx, ix, g, tag):
    for t in (0.0, 75.0):
        o = orc.score(hx, g, AND, t)
        for form, fname in ((issl.FORM_ORDERED, "ordered"), (issl.FORM_PLAIN, "plain")):
            print(f"[{tag}] {fname} vs oracle, {g.size} guides, threshold {t}", flush=True)
            r = ix.score(g, method=AND, threshold=t, form=form)
            assert (bits(r.mit) == bits(o["mit"])).all(), (fname, t)
            assert (bits(r.cfd) == bits(o["cfd"])).all(), (fname, t)
            assert (bits(r.mit_sum) == bits(o["mit_sum"])).all(), (fname, t)
            assert (bits(r.cfd_sum) == bits(o["cfd_sum"])).all(), (fname, t)
            assert (r.occ_count == o["occ_count"]).all(), (fname, t)
            assert (r.site_count == o["site_count"]).all(), (fname, t)
            assert (r.passed == o["pass"]).all(), (fname, t)


def batched_vs_ordered(ix, g, tag):
    """Every guide, both site-word layouts forced, both precisions."""
    print(f"[{tag}] ordered, {g.size} guides", flush=True)
    ref = ix.score(g, method=AND, threshold=0.0, form=issl.FORM_ORDERED)
    assert (ref.site_count > 0).mean() > 0.5
    built, used, _ = ix.residual()
    assert built
    try:
        for res in (True, False):
            ix.set_residual(res)
            for prec, rtol in ((issl.PREC_FP64, 1e-12), (issl.PREC_FP32C, 1e-6)):
                print(f"[{tag}] batched residual={res} precision={prec}", flush=True)
                r = ix.score(g, method=AND, threshold=0.0, form=issl.FORM_BATCHED, precision=prec)
                assert (r.site_count == ref.site_count).all(), (res, prec)
                assert (r.occ_count == ref.occ_count).all(), (res, prec)
                for f in ("mit_sum", "cfd_sum", "mit", "cfd"):
                    np.testing.assert_allclose(getattr(r, f), getattr(ref, f), rtol=rtol, atol=0,
                                               err_msg=f"{tag} {f} residual={res} prec={prec}")
    finally:
        ix.set_residual(used)


# ---------------------------------------------------------------- 100,000 guides

@pytest.fixture(scope="module")
def c2(c1_index):
    hx, ix, _, _ = c1_index
    print("[gpu-c2] generating the 100,000-guide mix", flush=True)
    sigs, _, masks, guides, kinds = workloads.make("genome100m_10x10_q100k", threads=threads())
    assert np.array_equal(sigs, hx.sigs) and np.array_equal(masks, hx.masks)
    return hx, ix, guides, kinds


def test_c2_ordered_bit_exact_vs_oracle(c2):
    hx, ix, guides, kinds = c2
    exact_vs_oracle(hx, ix, guides[sample(kinds, 48, 21)], "gpu-c2")


def test_c2_batched_vs_ordered_both_layouts(c2):
    hx, ix, guides, kinds = c2
    batched_vs_ordered(ix, guides, "gpu-c2")


# ---------------------------------------------------------------- 300 M sites

@pytest.fixture(scope="module")
def c3(gpu_available):
    print(f"[gpu-c3] building genome300m_10x10 ({threads()} threads)", flush=True)
    sigs, occ, masks, guides, kinds = workloads.make("genome300m_10x10", threads=threads())
    hx = issl.build_lists(sigs, occ, masks, threads=threads())
    del sigs, occ
    print("[gpu-c3] uploading", flush=True)
    ix = issl.Index.from_arrays(hx)
    assert hx.n_sites * hx.n_slices > (1 << 31)
    yield hx, ix, guides, kinds
    ix.close()


def test_c3_exact_forms_bit_exact_vs_oracle(c3):
    hx, ix, guides, kinds = c3
    exact_vs_oracle(hx, ix, guides[sample(kinds, 32, 31)], "gpu-c3")


def test_c3_batched_vs_ordered_both_layouts(c3):
    hx, ix, guides, kinds = c3
    batched_vs_ordered(ix, guides, "gpu-c3")


def test_c3_threshold_calls(c3):
    hx, ix, guides, kinds = c3
    print("[gpu-c3] threshold 75 calls", flush=True)
    o = ix.score(guides, method=AND, threshold=75.0, form=issl.FORM_ORDERED)
    full = ix.score(guides, method=AND, threshold=0.0, form=issl.FORM_ORDERED)
    cut = (10000.0 - 7500.0) / 75.0
    near = (np.abs(full.mit_sum - cut) <= 1e-6 * cut) | (np.abs(full.cfd_sum - cut) <= 1e-6 * cut)
    for prec in (issl.PREC_FP64, issl.PREC_FP32C):
        r = ix.score(guides, method=AND, threshold=75.0, form=issl.FORM_BATCHED, precision=prec)
        assert ((r.passed == o.passed) | near).all(), prec
    assert (o.passed == ord("0")).any() and (o.passed == ord("1")).any()


def test_c3_site_counts_vs_brute_force(c3):
    hx, ix, guides, kinds = c3
    idx = sample(kinds, 1, 33)
    r = ix.score(guides[idx], method=AND, threshold=0.0, form=issl.FORM_ORDERED)
    for j, i in enumerate(idx):
        print(f"[gpu-c3] brute force over 300 M sites, guide {i}", flush=True)
        ids = orc.candidates_brute(hx, int(guides[i]))
        assert r.site_count[j] == ids.size


# ---------------------------------------------------------------- both resident, served

def test_c4_server_two_indexes_bit_identical(c1_index, c3):
    hx1, ix1, g1, _ = c1_index
    hx3, ix3, g3, _ = c3
    direct = {0: ix1, 1: ix3}
    pools = {0: g1, 1: g3}
    rng = np.random.default_rng(0xC4)
    print("[gpu-c4] admitting both indexes into one server", flush=True)
    with issl.Server(max_batch=16384, slots=3) as srv:
        ids = {0: srv.add_index(hx1), 1: srv.add_index(hx3)}
        for ix in direct.values():
            ix.set_tuning()   # the server's indexes keep the default launch shape
        jobs = []
        for k in range(200):
            which = int(rng.integers(0, 2))
            n = max(1, int(np.exp(rng.uniform(0.0, np.log(20000.0)))))
            lo = int(rng.integers(0, pools[which].shape[0] - 1))
            g = np.resize(pools[which][lo:], n) if lo + n > pools[which].shape[0] else pools[which][lo:lo + n]
            kw = [dict(form=issl.FORM_BATCHED, precision=issl.PREC_FP32C),
                  dict(form=issl.FORM_BATCHED, precision=issl.PREC_FP64),
                  dict(form=issl.FORM_ORDERED, precision=issl.PREC_FP64)][k % 3]
            t = (0.0, 75.0)[(k // 3) % 2]
            jobs.append((which, g, t, kw, srv.submit(ids[which], g, threshold=t, **kw)))
        print("[gpu-c4] waiting for 200 requests", flush=True)
        for which, g, t, kw, req in jobs:
            got = req.wait(timeout=300.0)
            want = direct[which].score(g, threshold=t, **kw)
            assert same_bits(got, want), (which, g.size, t, kw)
        st = srv.stats()
        assert st.requests_failed == 0 and st.requests_done == 200
        assert st.peak_inflight >= 2


# ---------------------------------------------------------------- heavy tail

@pytest.fixture(scope="module")
def heavy(gpu_available):
    print(f"[gpu-heavy] building genome100m_10x10_heavy ({threads()} threads)", flush=True)
    sigs, occ, masks, guides, kinds = workloads.make("genome100m_10x10_heavy", threads=threads())
    hx = issl.build_lists(sigs, occ, masks, threads=threads())
    del sigs, occ
    print("[gpu-heavy] uploading", flush=True)
    ix = issl.Index.from_arrays(hx)
    yield hx, ix, guides, kinds
    ix.close()


def test_heavy_tail_is_heavy(heavy):
    """The repeat-neighbourhood guides visit 100-1000x the median posting volume."""
    hx, ix, guides, kinds = heavy
    P = np.zeros(guides.shape[0], dtype=np.int64)
    base = 0
    for m in hx.masks.tolist():
        pos = [j for j in range(20) if (m >> j) & 1]
        key = np.zeros(guides.shape[0], dtype=np.int64)
        for j, p in enumerate(pos):
            key |= ((guides >> np.uint64(2 * p)) & np.uint64(3)).astype(np.int64) << (2 * j)
        P += hx.list_sizes[base + key].astype(np.int64)
        base += 1 << (2 * len(pos))
    med = np.median(P)
    ratio = np.median(P[kinds == 3]) / med
    print(f"[gpu-heavy] repeat guides at {ratio:.0f}x the median postings (max {P.max() / med:.0f}x)")
    assert 100 <= ratio <= 1000


def test_heavy_exact_forms_vs_oracle(heavy):
    hx, ix, guides, kinds = heavy
    idx = np.sort(np.concatenate([np.nonzero(kinds == 3)[0][:8], np.nonzero(kinds == 1)[0][:24]]))
    exact_vs_oracle(hx, ix, guides[idx], "gpu-heavy")


def test_heavy_batched_vs_ordered(heavy):
    hx, ix, guides, kinds = heavy
    batched_vs_ordered(ix, guides, "gpu-heavy")


def split_vs_unsplit(ix, g, split_at, tag):
    """ORDERED with long guides split over tiles against one tile per guide:
    the same bits for every guide, threshold off (split) and on (where the
    form keeps one tile per guide)."""
    try:
        for t in (0.0, 75.0):
            ix.set_tuning(chunk=4096, depth=8, grid_per_cu=8, ord_depth=8, ord_grid=6,
                          ord_split=issl.SPLIT_NEVER)
            print(f"[{tag}] ordered, one tile per guide, threshold {t}", flush=True)
            one = ix.score(g, method=AND, threshold=t, form=issl.FORM_ORDERED)
            ix.set_tuning(chunk=4096, depth=8, grid_per_cu=8, ord_depth=8, ord_grid=6,
                          ord_split=split_at)
            print(f"[{tag}] ordered, split above {split_at} postings, threshold {t}", flush=True)
            sp = ix.score(g, method=AND, threshold=t, form=issl.FORM_ORDERED)
            assert same_bits(sp, one), (tag, t)
    finally:
        ix.set_tuning()


def test_heavy_ordered_split_bits(heavy):
    """The heavy guides (over 100,000 postings) split, the rest one tile."""
    hx, ix, guides, kinds = heavy
    split_vs_unsplit(ix, guides, 65536, "gpu-heavy")


def test_c2_ordered_split_bits(c2):
    """Every C2 guide over 1,024 postings asks for term entries: the 100,000
    guides need more than a launch holds, so some are split and the rest fall
    back to one tile."""
    hx, ix, guides, kinds = c2
    split_vs_unsplit(ix, guides, 1024, "gpu-c2")
