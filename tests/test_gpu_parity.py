"""HIP scorer vs the reference's golden vectors and the CPU oracle (MI355X).

Tolerances:
  * PLAIN and ORDERED (FP64, in-order accumulation): bit-exact scores, sums,
    counts and calls for every method and every golden threshold (0, 75,
    100, 99.99, 1, -5).
  * BATCHED FP64: scores within 1e-12 relative (summation order differs);
    counts and candidate sets exact.
  * BATCHED FP32C (FP32 products, compensated FP32 lane sums, FP64
    combination of lanes and tiles): scores within
    1e-6 relative with the threshold disabled; calls identical with the
    threshold on except guides whose FP64 sum lies within 1e-6 relative of
    the cutoff.
  * Threshold on, BATCHED: the exit hit, counts and calls equal the
    reference's for every guide whose running sums never come within the
    form's tolerance (1e-12 FP64, 1e-6 FP32C) of the cut (the oracle's exit
    margin) and whose scores are not that close to the threshold.
"""
import itertools

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import GOLDEN_CASES, GOLDEN_THRESHOLDS, call_margin, golden_index, load_golden, tkey

pytestmark = pytest.mark.gpu

METHODS = range(5)
THRESHOLDS = (0, 75)


@pytest.fixture(scope="module")
def cases(gpu_available):
    out = {}
    for name in GOLDEN_CASES:
        print(f"[gpu-parity] opening {name}", flush=True)
        hx = golden_index(name)
        out[name] = (hx, issl.Index.from_arrays(hx), load_golden(name))
    yield out
    for _, ix, _ in out.values():
        ix.close()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


EXACT_FORMS = {"plain": issl.FORM_PLAIN, "ordered": issl.FORM_ORDERED}


@pytest.mark.parametrize("form", sorted(EXACT_FORMS))
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_plain_bit_exact_vs_reference(cases, name, form):
    hx, ix, gd = cases[name]
    g = gd["guides"]
    for m, t in itertools.product(METHODS, GOLDEN_THRESHOLDS):
        print(f"[gpu-parity] {form} {name} m{m} t{t}", flush=True)
        r = ix.score(g, method=m, threshold=t, form=EXACT_FORMS[form])
        key = f"m{m}_t{tkey(t)}"
        assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), key
        assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), key
        assert (r.passed == gd[key + "_pass"]).all(), key


@pytest.mark.parametrize("form", sorted(EXACT_FORMS))
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_plain_sums_and_counts_vs_oracle(cases, name, form):
    hx, ix, gd = cases[name]
    g = gd["guides"]
    for m, t in itertools.product(METHODS, GOLDEN_THRESHOLDS):
        o = orc.score(hx, g, m, t)
        r = ix.score(g, method=m, threshold=t, form=EXACT_FORMS[form])
        assert (bits(r.mit_sum) == bits(o["mit_sum"])).all()
        assert (bits(r.cfd_sum) == bits(o["cfd_sum"])).all()
        assert (r.occ_count == o["occ_count"]).all()
        assert (r.site_count == o["site_count"]).all()


@pytest.mark.parametrize("precision,rtol", [(issl.PREC_FP64, 1e-12), (issl.PREC_FP32C, 1e-6)])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_batched_vs_reference(cases, name, precision, rtol):
    hx, ix, gd = cases[name]
    g = gd["guides"]
    for m in METHODS:
        print(f"[gpu-parity] batched p{precision} {name} m{m}", flush=True)
        # threshold disabled: scores within tolerance, counts exact
        r = ix.score(g, method=m, threshold=0.0, form=issl.FORM_BATCHED, precision=precision)
        key = f"m{m}_t0"
        for f in ("mit", "cfd"):
            want = gd[key + "_" + f]
            np.testing.assert_allclose(r.__dict__[f], want, rtol=rtol, atol=0)
        o = orc.score(hx, g, m, 0.0)
        assert (r.site_count == o["site_count"]).all()
        assert (r.occ_count == o["occ_count"]).all()
        # threshold on: identical calls away from the cutoff
        for t in GOLDEN_THRESHOLDS[1:]:
            rt = ix.score(g, method=m, threshold=t, form=issl.FORM_BATCHED, precision=precision)
            want = gd[f"m{m}_t{tkey(t)}_pass"]
            ot = orc.score(hx, g, m, t, margin=True)
            near = (ot["margin"] <= rtol) | (call_margin(ot, m, t) <= rtol)
            assert (rt.passed[~near] == want[~near]).all(), (m, t)


@pytest.mark.parametrize("precision,tol", [(issl.PREC_FP64, 1e-12), (issl.PREC_FP32C, 1e-6)])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_batched_exit_point_matches_reference(cases, name, precision, tol):
    """Threshold on: the finalizer's rescan stops at the reference's exit hit.
    Tile sums regroup the additions, so a guide may stop one hit earlier or
    later only where the reference's running sums come within rounding of the
    cut (the oracle's exit margin); every other guide must stop at the same
    hit, with the same counts, sums within tolerance and the same call."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    for m, t in itertools.product(METHODS, GOLDEN_THRESHOLDS[1:]):
        o = orc.score(hx, g, m, t, margin=True)
        r = ix.score(g, method=m, threshold=t, form=issl.FORM_BATCHED, precision=precision)
        free = (o["margin"] > tol) & (call_margin(o, m, t) > tol)
        what = f"{name} m{m} t{t}"
        assert (~free).sum() <= max(2, len(g) // 100), what
        assert (r.site_count[free] == o["site_count"][free]).all(), what
        assert (r.occ_count[free] == o["occ_count"][free]).all(), what
        np.testing.assert_allclose(r.mit_sum[free], o["mit_sum"][free], rtol=tol, atol=0,
                                   err_msg=what)
        np.testing.assert_allclose(r.cfd_sum[free], o["cfd_sum"][free], rtol=tol, atol=0,
                                   err_msg=what)
        assert (r.passed[free] == o["pass"][free]).all(), what


@pytest.mark.parametrize("shape", [dict(depth=4, grid_per_cu=2), dict(depth=8, grid_per_cu=6),
                                   dict(depth=16, grid_per_cu=8)])
@pytest.mark.parametrize("name", ["tiny4k_5x4", "mid50k_10x10", "small200k_8x8", "uniform1m_5x4"])
def test_ordered_shapes_and_layouts(cases, name, shape):
    """ORDERED under every load depth and both site-word layouts: the
    reference's bits at every golden threshold, with distinct sites marked up
    to the exit."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    built, used, _ = ix.residual()
    try:
        ix.set_tuning(chunk=1024, heavy_kpost=0, **shape)
        for res in ([False, True] if built else [False]):
            ix.set_residual(res)
            for m, t in itertools.product((issl.METHOD_MIT_AND_CFD, issl.METHOD_AVG_MIT_CFD,
                                           issl.METHOD_CFD), GOLDEN_THRESHOLDS):
                r = ix.score(g, method=m, threshold=t, form=issl.FORM_ORDERED)
                key = f"m{m}_t{tkey(t)}"
                what = f"{name} {shape} res {res} {key}"
                assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), what
                assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), what
                assert (r.passed == gd[key + "_pass"]).all(), what
            o = orc.score(hx, g, 2, 75.0, distinct=True)
            ix.distinct_reset()
            ix.score(g, method=2, threshold=75.0, form=issl.FORM_ORDERED, flags=issl.FLAG_DISTINCT)
            assert ix.distinct_count() == sum(bin(int(w)).count("1") for w in o["distinct"])
    finally:
        ix.set_residual(used)
        ix.set_tuning()


@pytest.mark.parametrize("name", ["tiny4k_5x4", "mid50k_10x10", "small200k_8x8", "uniform1m_5x4"])
def test_ordered_split(cases, name):
    """ORDERED with the split threshold at 64 postings: with the exit off every
    longer guide is split over 64-posting tiles that run in parallel and write
    their hits' terms, which one wave adds in order; with the exit armed the
    guides stay one tile each.  The reference's bits at every golden
    threshold, both site-word layouts, counts and distinct sites."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    built, used, _ = ix.residual()
    try:
        ix.set_tuning(chunk=64, depth=8, grid_per_cu=4, heavy_kpost=0, ord_split=64)
        assert ix.tuning().ord_split == 64
        for res in ([False, True] if built else [False]):
            ix.set_residual(res)
            for m, t in itertools.product(METHODS, GOLDEN_THRESHOLDS):
                r = ix.score(g, method=m, threshold=t, form=issl.FORM_ORDERED)
                key = f"m{m}_t{tkey(t)}"
                what = f"{name} split res {res} {key}"
                assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), what
                assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), what
                assert (r.passed == gd[key + "_pass"]).all(), what
            o = orc.score(hx, g, 2, 75.0, distinct=True)
            for t in (75.0, 0.0):
                ot = o if t == 75.0 else orc.score(hx, g, 2, 0.0, distinct=True)
                r = ix.score(g, method=2, threshold=t, form=issl.FORM_ORDERED)
                assert (r.occ_count == ot["occ_count"]).all() and (r.site_count == ot["site_count"]).all()
                ix.distinct_reset()
                ix.score(g, method=2, threshold=t, form=issl.FORM_ORDERED, flags=issl.FLAG_DISTINCT)
                assert ix.distinct_count() == sum(bin(int(w)).count("1") for w in ot["distinct"]), t
    finally:
        ix.set_residual(used)
        ix.set_tuning()


BATCHED_MODES = {
    # every guide through the persistent tile kernel, three tile shapes
    "tiles": dict(chunk=1024, depth=4, grid_per_cu=4, heavy_kpost=0),
    "tiles_short": dict(chunk=64, depth=16, grid_per_cu=1, heavy_kpost=0),
    "tiles_deep": dict(chunk=512, depth=16, grid_per_cu=8, heavy_kpost=0),
    # one wave per guide only
    "direct": dict(direct_depth=8, heavy_kpost=issl._capi.HEAVY_NEVER),
    # guides with more than 1,024 postings go to the tile kernel
    "hybrid": dict(chunk=512, depth=4, grid_per_cu=2, direct_depth=4, heavy_kpost=1),
}


@pytest.mark.parametrize("mode", sorted(BATCHED_MODES))
@pytest.mark.parametrize("name", ["tiny4k_5x4", "small200k_8x8", "uniform1m_5x4"])
def test_batched_modes(cases, name, mode):
    """Each scheduling mode of the BATCHED form against the reference's goldens."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    ix.set_tuning(**BATCHED_MODES[mode])
    try:
        for m in (issl.METHOD_MIT_AND_CFD, issl.METHOD_AVG_MIT_CFD):
            for t in THRESHOLDS:
                r = ix.score(g, method=m, threshold=float(t), form=issl.FORM_BATCHED,
                             precision=issl.PREC_FP64, flags=issl.FLAG_DISTINCT)
                key = f"m{m}_t{t}"
                if mode == "direct":
                    # one wave adds a guide's hits in the reference's order
                    assert (bits(r.mit) == bits(gd[key + "_mit"])).all(), key
                    assert (bits(r.cfd) == bits(gd[key + "_cfd"])).all(), key
                else:
                    np.testing.assert_allclose(r.mit, gd[key + "_mit"], rtol=1e-12)
                    np.testing.assert_allclose(r.cfd, gd[key + "_cfd"], rtol=1e-12)
                assert (r.passed == gd[key + "_pass"]).all(), key
            o = orc.score(hx, g, m, 0.0)
            r = ix.score(g, method=m, form=issl.FORM_BATCHED, precision=issl.PREC_FP32C)
            assert (r.site_count == o["site_count"]).all()
            assert (r.occ_count == o["occ_count"]).all()
            np.testing.assert_allclose(r.cfd, o["cfd"], rtol=1e-6)
    finally:
        ix.set_tuning()


@pytest.mark.parametrize("name", ["legacy39_5x4", "tiny4k_5x4", "mid50k_10x10"])
def test_candidate_sets_match_brute_force(cases, name):
    hx, ix, gd = cases[name]
    g = gd["guides"][:200]
    off, ids = ix.candidates(g)
    for i, q in enumerate(g):
        got = ids[off[i]:off[i + 1]]
        want = orc.candidates_brute(hx, int(q))
        assert np.array_equal(np.sort(got), want), i
        assert len(np.unique(got)) == len(got)


@pytest.mark.parametrize("form", [issl.FORM_PLAIN, issl.FORM_BATCHED, issl.FORM_ORDERED])
@pytest.mark.parametrize("t", THRESHOLDS)
def test_distinct_sites_match_oracle(cases, form, t):
    hx, ix, gd = cases["small200k_8x8"]
    g = gd["guides"]
    o = orc.score(hx, g, 2, float(t), distinct=True)
    want = sum(bin(int(w)).count("1") for w in o["distinct"])
    ix.distinct_reset()
    ix.score(g, method=2, threshold=float(t), form=form, flags=issl.FLAG_DISTINCT)
    assert ix.distinct_count() == want


def test_batch_composition_does_not_change_results(cases):
    hx, ix, gd = cases["mid50k_10x10"]
    g = gd["guides"]
    for prec in (issl.PREC_FP64, issl.PREC_FP32C):
        full = ix.score(g, form=issl.FORM_BATCHED, precision=prec)
        for lo, hi in [(0, 1), (17, 80), (300, 500)]:
            part = ix.score(g[lo:hi], form=issl.FORM_BATCHED, precision=prec)
            assert (bits(part.mit) == bits(full.mit[lo:hi])).all()
            assert (bits(part.cfd) == bits(full.cfd[lo:hi])).all()
        shuffled = np.random.default_rng(5).permutation(len(g))
        sh = ix.score(g[shuffled], form=issl.FORM_BATCHED, precision=prec)
        assert (bits(sh.mit) == bits(full.mit[shuffled])).all()


def test_validate_index(cases):
    hx, ix, gd = cases["tiny4k_5x4"]
    assert ix.validate() == 0
    bad = issl.IndexArrays(hx.sigs.copy(), hx.masks.copy(), hx.list_sizes.copy(),
                           hx.entries.copy(), hx.seq_len)
    # move one entry's id to a different site: wrong list and a duplicate id
    old_id = int(bad.entries[5]) & 0xFFFFFFFF
    bad.entries[5] = (bad.entries[5] & np.uint64(0xFFFFFFFF00000000)) | np.uint64(
        (old_id + 1) % hx.n_sites)
    with issl.Index.from_arrays(bad) as b:
        assert b.validate() > 0


def test_argument_errors(cases):
    hx, ix, gd = cases["tiny4k_5x4"]
    g = gd["guides"][:10]
    with pytest.raises(issl.IsslError):
        ix.score(g, max_dist=3)
    with pytest.raises(issl.IsslError):
        ix.score(g, method=7)
    with pytest.raises(issl.IsslError):
        ix.score(g, form=issl.FORM_PLAIN, precision=issl.PREC_FP32C)
    with pytest.raises(issl.IsslError):
        ix.score(g, form=issl.FORM_ORDERED, precision=issl.PREC_FP32C)
    with pytest.raises(issl.IsslError):
        ix.score(g, form=3)
    r = ix.score(g[:0])
    assert r.mit.shape == (0,)


def test_index_info_and_open_file(cases, tmp_path):
    hx, ix, gd = cases["tiny4k_5x4"]
    info = ix.info()
    assert info.n_sites == hx.n_sites and info.n_slices == 5 and info.max_width == 4
    p = str(tmp_path / "t.issl")
    issl.write_index(p, hx)
    with issl.Index.open(p) as f:
        a = f.score(gd["guides"])
        b = ix.score(gd["guides"])
        assert (bits(a.cfd) == bits(b.cfd)).all()


def test_autotune_keeps_results(cases, tmp_path):
    hx, ix, gd = cases["small200k_8x8"]
    g = gd["guides"]
    before = ix.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
    t = ix.autotune(g, str(tmp_path / "tune.txt"))
    assert t.chunk % 64 == 0 and t.us_per_guide > 0
    after = ix.score(g, form=issl.FORM_BATCHED, precision=issl.PREC_FP64)
    np.testing.assert_allclose(after.cfd, before.cfd, rtol=1e-12)
    assert (after.site_count == before.site_count).all()
    t2 = ix.autotune(g, str(tmp_path / "tune.txt"))  # from the cache
    assert (t2.chunk, t2.depth, t2.grid_per_cu) == (t.chunk, t.depth, t.grid_per_cu)
    # The key holds the bound on a guide's postings, so that indexes of one
    # site count and slice shape with different list lengths do not share a
    # tuning (the in-order split is only tried where guides can pass it).
    key = (tmp_path / "tune.txt").read_text().split()[0]
    assert key.startswith("t9/") and "/pm" in key, key


@pytest.mark.parametrize("form", [issl.FORM_PLAIN, issl.FORM_BATCHED, issl.FORM_ORDERED])
def test_timing_repeat_keeps_results(cases, form):
    """Timed launches (one or several dispatches per event pair) give the same results."""
    hx, ix, gd = cases["mid50k_10x10"]
    g = gd["guides"]
    ref = ix.score(g, form=form)
    for repeat in (1, 3):
        ix.timing_enable(True, repeat=repeat)
        r = ix.score(g, form=form)
        r = ix.score(g, form=form)
        ms, n = ix.timing_read()
        ix.timing_enable(False)
        assert n == 2 * repeat and ms > 0
        assert (bits(r.cfd) == bits(ref.cfd)).all()
        assert (bits(r.mit) == bits(ref.mit)).all()
        assert (r.site_count == ref.site_count).all()


@pytest.mark.parametrize("name", ["mid50k_10x10", "small200k_8x8", "uniform1m_5x4"])
def test_one_tile_launches(cases, name):
    """A chunk that covers the longest candidate sequence gives every guide one
    tile (finished by the wave that scores it), short chunks give many; the
    shapes and the early exit alternate, so each launch sequence must leave
    the tile queue's counters usable for the next."""
    hx, ix, gd = cases[name]
    g = gd["guides"]
    o0 = orc.score(hx, g, issl.METHOD_MIT_AND_CFD, 0.0)
    o75 = orc.score(hx, g, issl.METHOD_MIT_AND_CFD, 75.0)
    try:
        for step, (chunk, t) in enumerate([(1 << 20, 0.0), (1 << 20, 0.0), (1024, 0.0), (1 << 20, 0.0),
                                           (1 << 20, 75.0), (1 << 20, 0.0), (64, 0.0), (1 << 20, 0.0)]):
            ix.set_tuning(chunk=chunk, depth=8, grid_per_cu=4, heavy_kpost=0)
            r = ix.score(g, method=issl.METHOD_MIT_AND_CFD, threshold=t, form=issl.FORM_BATCHED,
                         precision=issl.PREC_FP32C)
            o = o0 if t == 0.0 else o75
            what = f"{name} step {step} chunk {chunk} t {t}"
            if t == 0.0:
                assert (r.site_count == o["site_count"]).all(), what
                assert (r.occ_count == o["occ_count"]).all(), what
                np.testing.assert_allclose(r.cfd, o["cfd"], rtol=1e-6, err_msg=what)
                np.testing.assert_allclose(r.mit, o["mit"], rtol=1e-6, err_msg=what)
            key = f"m{issl.METHOD_MIT_AND_CFD}_t{int(t)}"
            assert (r.passed == gd[key + "_pass"]).all(), what
    finally:
        ix.set_tuning()


@pytest.mark.parametrize("form", [issl.FORM_ORDERED, issl.FORM_BATCHED, issl.FORM_PLAIN])
def test_score_with_a_subset_of_columns(cases, form):
    """issl_score writes and copies back only the result columns the caller
    asks for (packed in the workspace); each subset equals the same columns
    of a full call: the pipeline module's (MIT, CFD, call), the CLI's (MIT,
    CFD, occurrences), single columns and none."""
    import ctypes
    from cracklingplusplus_gpu_amd import _capi
    print(f"[gpu-parity] column subsets form {form}", flush=True)
    hx, ix, gold = cases["mid50k_10x10"]
    g = np.ascontiguousarray(gold["guides"], dtype=np.uint64)
    n = g.shape[0]
    kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=75.0, form=form,
              precision=issl.PREC_FP64)
    full = ix.score(g, **kw)
    cols = {"mit": (np.float64, full.mit), "cfd": (np.float64, full.cfd),
            "occ_count": (np.uint64, full.occ_count), "site_count": (np.uint64, full.site_count),
            "pass": (np.uint8, full.passed), "mit_sum": (np.float64, full.mit_sum),
            "cfd_sum": (np.float64, full.cfd_sum)}
    p = ix.params(kw["method"], kw["threshold"], form, kw["precision"])
    for subset in (("mit", "cfd", "pass"), ("mit", "cfd", "occ_count"), ("pass",),
                   ("site_count", "cfd_sum"), ()):
        bufs = {k: np.full(n, 0xAB, dtype=cols[k][0]) for k in subset}
        r = _capi.issl_results(*[bufs[k].ctypes.data_as(ctypes.c_void_p) if k in bufs else None
                                 for k in ("mit", "cfd", "occ_count", "site_count", "pass",
                                           "mit_sum", "cfd_sum")])
        issl.check(issl.lib().issl_score(ix.handle, g.ctypes.data_as(ctypes.c_void_p), n,
                                         ctypes.byref(p), ctypes.byref(r)))
        for k in subset:
            want = cols[k][1]
            if want.dtype == np.float64:
                assert (bits(bufs[k]) == bits(want)).all(), (subset, k)
            else:
                assert (bufs[k].view(np.uint8) == np.ascontiguousarray(want).view(np.uint8)).all(), (subset, k)


@pytest.mark.parametrize("form", [issl.FORM_ORDERED, issl.FORM_BATCHED, issl.FORM_PLAIN])
def test_host_call_in_pieces(cases, form):
    """A host-buffer call of more than 262,144 guides runs in pieces whose
    staging and result copies overlap the device's work: 600,000 guides
    (the golden guides repeated, three pieces) give, guide by guide, the bits
    of a call on the golden guides alone; with the module's three columns
    as well as all seven."""
    print(f"[gpu-parity] host call in pieces, form {form}", flush=True)
    hx, ix, gold = cases["mid50k_10x10"]
    g = np.ascontiguousarray(gold["guides"], dtype=np.uint64)
    m = g.shape[0]
    big = np.ascontiguousarray(np.resize(g, 600_000))
    idx = np.arange(big.shape[0]) % m
    for t in (0.0, 75.0):
        kw = dict(method=issl.METHOD_MIT_AND_CFD, threshold=t, form=form, precision=issl.PREC_FP64)
        one = ix.score(g, **kw)
        many = ix.score(big, **kw)
        for f in ("mit", "cfd", "mit_sum", "cfd_sum"):
            assert (bits(getattr(many, f)) == bits(getattr(one, f))[idx]).all(), (t, f)
        for f in ("occ_count", "site_count", "passed"):
            assert (getattr(many, f) == getattr(one, f)[idx]).all(), (t, f)
