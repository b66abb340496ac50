"""Host side of libissl_gpu.so: ABI exports, tables, index format, generators (no GPU)."""
import hashlib
import os
import re
import tempfile

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import GOLDEN, ROOT, golden_index, load_golden
from cracklingplusplus_gpu_amd import _capi, workloads


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(issl_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = issl.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if getattr(lib, n, None) is None]
    assert not missing, missing
    # and the Python binding knows every one of them
    bound = {s[0] for s in _capi.SIGNATURES}
    assert names <= bound, sorted(names - bound)


def test_product_tables_equal_oracle_and_reference_fixture():
    k, v, p, a = issl.score_tables()
    ok_, ov = orc.mit_table()
    assert (k == ok_).all() and (v.view(np.uint64) == ov.view(np.uint64)).all()
    op, oa = orc.cfd_tables()
    assert (p.view(np.uint64) == op.view(np.uint64)).all()
    assert (a.view(np.uint64) == oa.view(np.uint64)).all()
    raw = np.fromfile(os.path.join(GOLDEN, "ref_test_genome_offTargets_indexed.issl"),
                      dtype=np.uint64)
    tab = raw[6:6 + 2 * 6195].reshape(-1, 2)
    assert (tab[:, 0] == k).all() and (tab[:, 1] == v.view(np.uint64)).all()


def test_mit_factored_form_equals_reference_table():
    """The tile kernels form each MIT value as ((t1 * t2) * t3) * 100 with t1 the
    product of (1 - weight) over the mismatch positions (1.0 for absent ones),
    t2 looked up by (count, last - first position) and t3 by count.  Computed
    here with IEEE doubles in that order, it must equal every entry of the
    reference fixture's table (ISSLCreateIndex.cpp:235-261 wrote it)."""
    w = [0.0, 0.0, 0.014, 0.0, 0.0, 0.395, 0.317, 0.0, 0.389, 0.079, 0.445, 0.508,
         0.613, 0.851, 0.732, 0.828, 0.615, 0.804, 0.685, 0.583]
    om = [1.0 - x for x in w] + [1.0] * 13
    raw = np.fromfile(os.path.join(GOLDEN, "ref_test_genome_offTargets_indexed.issl"),
                      dtype=np.uint64)
    tab = raw[6:6 + 2 * 6195].reshape(-1, 2)
    for key, bits in tab.tolist():
        pos = [j for j in range(20) if (key >> (2 * j)) & 1]
        k = len(pos)
        assert 1 <= k <= 4
        span = pos[-1] - pos[0]
        gap = 19.0 if k == 1 else float(span) / float(k - 1)
        t2 = 1.0 / ((19.0 - gap) / 19.0 * 4.0 + 1)
        t3 = 1.0 / float(k * k)
        p = pos + [32] * (4 - k)
        t1 = om[p[0]] * om[p[1]]
        t1 = t1 * om[p[2]]
        t1 = t1 * om[p[3]]
        v = ((t1 * t2) * t3) * 100.0
        assert np.float64(v).view(np.uint64) == bits, (pos, v)


def test_builder_reproduces_reference_fixture_index():
    """Sorted site text -> (sig, occ) -> slice lists equals the reference's test index."""
    legacy = golden_index("legacy39_5x4")
    txt = open(os.path.join(GOLDEN, "ref_test_genome_offTargets.fa"), "rb").read()
    sigs, occ = issl.count_sorted_sites(txt, 20, 21)
    assert sigs.shape == (39,) and int(occ.sum()) == 43
    assert (sigs == legacy.sigs).all()
    built = issl.build_lists(sigs, occ, legacy.masks)
    assert (built.list_sizes == legacy.list_sizes).all()
    assert (built.entries == legacy.entries).all()


def test_encode_matches_oracle():
    rng = np.random.default_rng(1)
    seqs = ["".join(rng.choice(list("ACGTN"), 20)) for _ in range(200)]
    got = issl.encode(seqs)
    want = [orc.lib().or_encode(s.encode()) for s in seqs]
    assert got.tolist() == want
    assert issl.decode(issl.encode(["ACGTACGTACGTACGTACGT"])) == ["ACGTACGTACGTACGTACGT"]


def test_index_roundtrip_and_file_hash(tmp_path):
    ix = issl.read_index(os.path.join(GOLDEN, "tiny4k_5x4.issl"))
    p = str(tmp_path / "x.issl")
    issl.write_index(p, ix)
    assert open(p, "rb").read() == open(os.path.join(GOLDEN, "tiny4k_5x4.issl"), "rb").read()
    back = issl.read_index(p)
    for f in ("sigs", "masks", "list_sizes", "entries"):
        assert (getattr(back, f) == getattr(ix, f)).all()


@pytest.mark.parametrize("name", ["mid50k_10x10", "small200k_8x8", "uniform1m_5x4"])
def test_generator_reproduces_golden_index(name, tmp_path):
    """The regenerated index is byte-identical to the one the goldens were made on."""
    gd = load_golden(name)
    sigs, occ, masks, guides, _ = workloads.make(name)
    assert (guides == gd["guides"]).all()
    p = str(tmp_path / "g.issl")
    issl.write_index(p, issl.build_lists(sigs, occ, masks))
    h = hashlib.sha256(open(p, "rb").read()).hexdigest()
    assert h == str(gd["issl_sha256"])


def test_synth_is_thread_count_invariant():
    a = workloads.make("small200k_8x8", threads=1)
    b = workloads.make("small200k_8x8", threads=7)
    for x, y in zip(a, b):
        assert (x == y).all()


def test_synth_sites_distinct_in_text_order():
    sigs, occ, masks, guides, kinds = workloads.make("mid50k_10x10")
    seqs = issl.decode(sigs[:2000])
    assert seqs == sorted(seqs)
    assert len(np.unique(sigs)) == sigs.shape[0]
    assert (occ >= 1).all()
    assert [bin(int(m)).count("1") for m in masks] == [10] * 10
    assert np.bincount(kinds, minlength=5).tolist() == [200, 0, 50, 150, 100]


def test_geometric_convention_failures():
    """occ = 1 + F with P(F=k) = (1-p)^k p: mean 1.25 at p = 0.8."""
    sigs, occ, _, _, _ = workloads.make("uniform1m_5x4")
    assert abs(occ.mean() - 1.25) < 0.01
    assert abs((occ == 1).mean() - 0.8) < 0.01


def test_parse_masks():
    text = "11110000000000000000\n00001111000000000000\r\n00000000000000001111\n"
    m = issl.parse_masks(text)
    assert m.tolist() == [0xF, 0xF0, 0xF0000]


def test_list_count_and_width_limit():
    assert issl.list_count(np.array([0xF, 0xFF], dtype=np.uint64)) == 256 + 65536
    assert issl.list_count(np.array([(1 << 17) - 1], dtype=np.uint64)) == 0


def test_read_rejects_bad_files(tmp_path):
    ix = issl.read_index(os.path.join(GOLDEN, "tiny4k_5x4.issl"))
    p = str(tmp_path / "t.issl")
    issl.write_index(p, ix)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-8])
    with pytest.raises(issl.IsslError) as e:
        issl.read_index(p)
    assert e.value.code == _capi.ISSL_ERR_FORMAT
    bad = bytearray(data)
    bad[0:8] = np.uint64(ix.n_sites + 1).tobytes()
    open(p, "wb").write(bytes(bad))
    with pytest.raises(issl.IsslError):
        issl.read_index(p)
    with pytest.raises(issl.IsslError) as e:
        issl.read_index(str(tmp_path / "missing.issl"))
    assert e.value.code == _capi.ISSL_ERR_IO


def test_count_sorted_sites_rejects_ragged_file():
    with pytest.raises(issl.IsslError):
        issl.count_sorted_sites(b"ACGTACGTACGTACGTACGT\nACG\n", 20, 21)


def test_scoring_calls_fail_loudly_without_gpu():
    if issl.device_count() > 0:
        pytest.skip("a GPU is present")
    ix = issl.read_index(os.path.join(GOLDEN, "tiny4k_5x4.issl"))
    with pytest.raises(issl.IsslError) as e:
        issl.Index.from_arrays(ix)
    assert e.value.code == _capi.ISSL_ERR_DEVICE


def _index21(tmp_path):
    """A 21-position index (ISSLCreateIndex accepts seqLength up to 21,
    ISSLCreateIndex.cpp:296), five contiguous 4-position slices."""
    rng = np.random.default_rng(21)
    sigs = np.unique(rng.integers(0, 1 << 42, size=3000, dtype=np.uint64))
    occ = np.ones(sigs.shape[0], dtype=np.uint32)
    masks = np.array([0xF << (4 * s) for s in range(5)], dtype=np.uint64)
    return issl.build_lists(sigs, occ, masks, seq_len=21), occ


def test_21_position_index_is_refused(tmp_path):
    """The reference takes the distance on the whole 64-bit XOR
    (lib/ISSLScoringModule.cpp:286-290, ISSLScoreOfftargets.cpp:360-366), so a
    21-position index scores differently from a 20-position distance; every
    loader refuses it with ISSL_ERR_FORMAT before looking for a device."""
    hx, occ = _index21(tmp_path)
    assert hx.seq_len == 21 and int(hx.sigs.max()) >= (1 << 40)
    p = str(tmp_path / "i21.issl")
    issl.write_index(p, hx)
    calls = [lambda: issl.Index.from_arrays(hx),
             lambda: issl.Index.open(p),
             lambda: issl.Index.from_arrays_streamed(hx),
             lambda: issl.Index.open_streamed(p),
             lambda: issl.Index.build(hx.sigs, occ, hx.masks, seq_len=21)]
    for call in calls:
        with pytest.raises(issl.IsslError) as e:
            call()
        assert e.value.code == _capi.ISSL_ERR_FORMAT, e.value
        assert "20" in str(e.value)


def test_wide_signatures_in_20_position_index_are_refused():
    ix = issl.read_index(os.path.join(GOLDEN, "tiny4k_5x4.issl"))
    sigs = ix.sigs.copy()
    sigs[7] |= np.uint64(1 << 41)
    bad = issl.IndexArrays(sigs, ix.masks, ix.list_sizes, ix.entries, 20)
    for call in (lambda: issl.Index.from_arrays(bad),
                 lambda: issl.Index.from_arrays_streamed(bad),
                 lambda: issl.Index.build(sigs, ix.occurrences(), ix.masks)):
        with pytest.raises(issl.IsslError) as e:
            call()
        assert e.value.code == _capi.ISSL_ERR_FORMAT, e.value
        assert "position 19" in str(e.value)
