"""Drop-in CLI and pipeline module: argument handling (CPU) and outputs (GPU).

The CLI mirrors ISSLScoreOfftargets/ISSLScoreOfftargets.cpp:36-647 and the
module lib/ISSLScoringModule.cpp:13-513; expected values come from the
reference's golden vectors and the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import GOLDEN, ROOT, golden_index, load_golden

BIN = os.path.join(ROOT, "cracklingplusplus_gpu_amd", "bin")
CLI = os.path.join(BIN, "ISSLScoreOfftargets")
DEMO = os.path.join(BIN, "issl_module_demo")
TINY = os.path.join(GOLDEN, "tiny4k_5x4.issl")
CLI_METHOD = {"and": 2, "or": 3, "avg": 4, "mit": 0, "cfd": 1}


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)


def test_cli_usage_and_bad_method():
    r = run([CLI])
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = run([CLI, TINY, "q.txt", "4", "75", "best"])
    assert r.returncode == 1 and "Invalid scoring method" in r.stderr


def test_module_rejects_bad_method(tmp_path):
    g = tmp_path / "g.txt"
    g.write_text("ACGTACGTACGTACGTACGTAGG 1 0 ?\n")
    r = run([DEMO, TINY, str(g), "best", "75", "ultralow", "1", "0"])
    assert r.returncode == 1 and "Scoring method is not valid." in r.stderr


def test_module_disabled_skips(tmp_path):
    g = tmp_path / "g.txt"
    g.write_text("ACGTACGTACGTACGTACGTAGG 1 0 ?\n")
    r = run([DEMO, TINY, str(g), "mit", "75", "ultralow", "0", "0"])
    assert r.returncode == 0
    assert "Skipping Off-target scoring" in r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT")][0].split("\t")
    assert line[4] == "?" and float(line[2]) == -1.7976931348623157e308


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["", "plain", "batched"])
@pytest.mark.parametrize("t", ["75", "0", "99.99", "-5"])
@pytest.mark.parametrize("name", ["and", "or", "avg", "mit", "cfd"])
def test_cli_output_matches_reference(tmp_path, name, t, form, gpu_available):
    """stdout and the results file, byte for byte (ISSLScoreOfftargets.cpp:
    590-625: SEQ, MIT and CFD as %f or -1, the occurrence-weighted count).
    The default (ORDERED) and PLAIN forms give the reference's sums bit for
    bit; BATCHED's FP64 sums differ from them by rounding only, below the six
    decimals printed for these guides."""
    from conftest import tkey
    gd = load_golden("tiny4k_5x4")
    hx = golden_index("tiny4k_5x4")
    g = gd["guides"]
    seqs = issl.decode(g)
    q = tmp_path / "queries.txt"
    q.write_text("".join(s + "\n" for s in seqs))
    out = tmp_path / "res.txt"
    env = dict(os.environ)
    env.pop("ISSL_FORM", None)
    if form:
        env["ISSL_FORM"] = form
    r = run([CLI, TINY, str(q), "4", t, name, str(out)], env=env)
    assert r.returncode == 0, r.stderr
    m = CLI_METHOD[name]
    key = f"m{m}_t{tkey(float(t))}"
    mit, cfd = gd[key + "_mit"], gd[key + "_cfd"]
    o = orc.score(hx, g, m, float(t), distinct=True)
    want = []
    for i in range(len(seqs)):
        ms = "%f" % mit[i] if m != 1 else "-1"
        cs = "%f" % cfd[i] if m != 0 else "-1"
        want.append((seqs[i], ms, cs, int(o["occ_count"][i])))
    lines = [l for l in r.stdout.splitlines() if l.count("\t") == 2]
    assert len(lines) == len(seqs)
    file_lines = out.read_text().splitlines()
    assert file_lines[0] == "# Query\tMIT_Score\tCFD_Score\tOfftarget_Count"
    assert len(file_lines) == len(seqs) + 1
    for i, (sq, ms, cs, cnt) in enumerate(want):
        assert lines[i] == f"{sq}\t{ms}\t{cs}", i
        assert file_lines[1 + i] == f"{sq}\t{ms}\t{cs}\t{cnt}", i
    want_distinct = sum(bin(int(w)).count("1") for w in o["distinct"])
    assert f"Distinct off-targets: {want_distinct}" in r.stderr
    assert f"Total off-targets scored: {int(o['occ_count'].sum())}" in r.stderr


CLI_MMF = os.path.join(BIN, "ISSLScoreOfftargetsMMF")


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [1, 0])
@pytest.mark.parametrize("name,t", [("and", "75"), ("avg", "0"), ("mit", "75"), ("cfd", "75")])
def test_mmf_cli_output_matches_reference(tmp_path, name, t, slots, gpu_available):
    """ISSLScoreOfftargetsMMF: the reference's stdout lines only, from the
    streamed index, equal to the golden scores."""
    gd = load_golden("tiny4k_5x4")
    hx = golden_index("tiny4k_5x4")
    seqs = issl.decode(gd["guides"])
    q = tmp_path / "queries.txt"
    q.write_text("".join(s + "\n" for s in seqs))
    env = dict(os.environ, ISSL_STREAM_BYTES=str(slots * 16 * hx.n_sites))
    r = run([CLI_MMF, TINY, str(q), "4", t, name], env=env)
    assert r.returncode == 0, r.stderr
    m = CLI_METHOD[name]
    key = f"m{m}_t{t}"
    want = []
    for i, s in enumerate(seqs):
        ms = "%f" % gd[key + "_mit"][i] if m != 1 else "-1"
        cs = "%f" % gd[key + "_cfd"][i] if m != 0 else "-1"
        want.append(f"{s}\t{ms}\t{cs}")
    assert r.stdout.splitlines() == want


def test_mmf_cli_usage_and_bad_method():
    r = run([CLI_MMF])
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = run([CLI_MMF, TINY, "q.txt", "4", "75", "best"])
    assert r.returncode == 1 and "Invalid scoring method" in r.stderr


@pytest.mark.gpu
def test_cli_rejects_ragged_query_file(tmp_path, gpu_available):
    q = tmp_path / "q.txt"
    q.write_text("ACGTACGTACGTACGTACGT\nACG\n")
    r = run([CLI, TINY, str(q), "4", "75", "and", str(tmp_path / "o.txt")])
    assert r.returncode == 1 and "not a multiple of the expected line length (21)" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("method,level", [("mitAndCfd", "ultralow"), ("avgMitCfd", "medium"),
                                          ("mit", "low"), ("cfd", "high"),
                                          ("mitOrCfd", "medium")])
def test_module_matches_reference(tmp_path, method, level, gpu_available):
    gd = load_golden("tiny4k_5x4")
    g = gd["guides"]
    seqs = issl.decode(g)
    rng = np.random.default_rng(3)
    uniq = rng.random(len(g)) < 0.8
    cons = rng.integers(0, 4, len(g))
    bt = rng.choice(["1", "0", "?"], len(g))
    f = tmp_path / "g.txt"
    f.write_text("".join(f"{s}AGG {int(u)} {c} {b}\n" for s, u, c, b in zip(seqs, uniq, cons, bt)))
    r = run([DEMO, TINY, str(f), method, "75", level, "1", "2"])
    assert r.returncode == 0, r.stderr
    m = issl.METHOD_NAMES[method]
    key = f"m{m}_t75"
    res = [l.split("\t") for l in r.stdout.splitlines() if l.startswith("RESULT")]
    assert len(res) == len(g)
    tested = 0
    for i, row in enumerate(res):
        keep = level == "ultralow" or (
            uniq[i] and (level == "low" or (cons[i] >= 2 and bt[i] != "0")))
        if not keep:
            assert row[4] == "?"
            continue
        tested += 1
        assert float(row[2]) == gd[key + "_mit"][i]
        assert float(row[3]) == gd[key + "_cfd"][i]
        assert row[4] == chr(gd[key + "_pass"][i])
    failed = sum(1 for i, row in enumerate(res) if row[4] == "0")
    assert f"{failed:,} of {tested:,} failed here." in r.stdout


REFTYPES_DEMO = os.path.join(ROOT, "oracle", "_ref", "issl_module_demo_reftypes")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REFTYPES_DEMO),
                    reason="built by `make -C oracle ref` where the reference sources are present")
@pytest.mark.parametrize("mmf", [False, True])
@pytest.mark.parametrize("method,level", [("mitAndCfd", "medium"), ("avgMitCfd", "ultralow")])
def test_module_built_on_reference_types_matches(tmp_path, method, level, mmf, gpu_available):
    """The drop-in module compiled against the pipeline's own include/util.hpp
    (cracklingConfig, guideResults, the CODE_* constants of lib/util.cpp)
    behaves exactly as the build on the mirror types."""
    gd = load_golden("tiny4k_5x4")
    seqs = issl.decode(gd["guides"])
    rng = np.random.default_rng(9)
    f = tmp_path / "g.txt"
    f.write_text("".join(f"{s}AGG {int(rng.random() < 0.8)} {rng.integers(0, 4)} "
                         f"{rng.choice(['1', '0', '?'])}\n" for s in seqs))
    args = [TINY, str(f), method, "75", level, "1", "2"] + (["mmf"] if mmf else [])
    a = run([DEMO] + args)
    b = run([REFTYPES_DEMO] + args)
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    assert a.stdout == b.stdout
    assert sum(1 for l in b.stdout.splitlines() if l.startswith("RESULT")) == len(seqs)


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [1, 2, 0])
@pytest.mark.parametrize("method,level", [("mitAndCfd", "medium"), ("cfd", "ultralow")])
def test_module_mmf_matches_resident(tmp_path, method, level, slots, gpu_available):
    """The memory-mapped module (the Crackling pipeline's choice when
    offTarget.memoryMapped is set) streams the index through HBM slice
    buffers and gives the resident module's results bit for bit."""
    hx = golden_index("tiny4k_5x4")
    gd = load_golden("tiny4k_5x4")
    seqs = issl.decode(gd["guides"])
    rng = np.random.default_rng(11)
    f = tmp_path / "g.txt"
    f.write_text("".join(f"{s}AGG {int(rng.random() < 0.8)} {rng.integers(0, 4)} "
                         f"{rng.choice(['1', '0', '?'])}\n" for s in seqs))
    args = [TINY, str(f), method, "75", level, "1", "2"]
    env = dict(os.environ, ISSL_STREAM_BYTES=str(slots * 16 * hx.n_sites))
    a = run([DEMO] + args)
    b = run([DEMO] + args + ["mmf"], env=env)
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    assert "Mapping ISSL Index." in b.stdout and "Loading ISSL Index." not in b.stdout
    ra = [l for l in a.stdout.splitlines() if l.startswith("RESULT") or "failed here" in l]
    rb = [l for l in b.stdout.splitlines() if l.startswith("RESULT") or "failed here" in l]
    assert ra == rb
    assert sum(1 for l in rb if l.startswith("RESULT")) == len(seqs)


EXTRACT = os.path.join(BIN, "ExtractOfftargets")


def test_extract_cli_usage():
    r = run([EXTRACT, "out.txt"])
    assert r.returncode == 1 and "Usage:" in r.stderr


# ---------------------------------------------------------------- index builder

CREATE = os.path.join(BIN, "ISSLCreateIndex")


def sites_text(hx, ending="\n"):
    """Sorted site lines as ExtractOfftargets writes them: each distinct site
    repeated by its occurrence count, ids in text order."""
    occ = hx.occurrences()
    seqs = issl.decode(hx.sigs)
    assert seqs == sorted(seqs)
    return "".join((sq[:hx.seq_len] + ending) * int(o) for sq, o in zip(seqs, occ))


def masks_text(masks, seq_len=20, ending="\n"):
    return "".join("".join("1" if (int(m) >> j) & 1 else "0" for j in range(seq_len)) + ending
                   for m in masks)


@pytest.mark.parametrize("ending", ["\n", "\r\n"])
def test_create_index_reproduces_committed_index(tmp_path, ending):
    """ISSLCreateIndex.cpp:284-493 on the committed index's sites: the same
    bytes, whatever the line ending (getLineEnding, :32-52)."""
    hx = golden_index("tiny4k_5x4")
    sites = tmp_path / "sites.txt"
    sites.write_bytes(sites_text(hx, ending).encode())
    sc = tmp_path / "slices.txt"
    sc.write_bytes(masks_text(hx.masks, 20, ending).encode())
    out = tmp_path / "out.issl"
    r = run([CREATE, str(sites), str(sc), "20", str(out)])
    assert r.returncode == 0, r.stderr
    assert f"Number of sequences: {int(hx.occurrences().sum())}" in r.stdout
    assert out.read_bytes() == open(TINY, "rb").read()


def test_create_index_reproduces_reference_fixture_lists(tmp_path):
    """The reference's own test index (tests/data/test_genome_offTargets_indexed.issl,
    legacy layout): the same signatures, masks, list sizes and entries."""
    hx = golden_index("legacy39_5x4")
    sites = tmp_path / "sites.txt"
    sites.write_text(sites_text(hx))
    sc = tmp_path / "slices.txt"
    sc.write_text(masks_text(hx.masks))
    out = tmp_path / "out.issl"
    r = run([CREATE, str(sites), str(sc), "20", str(out)])
    assert r.returncode == 0, r.stderr
    got = issl.read_index(str(out))
    for f in ("sigs", "masks", "list_sizes", "entries"):
        assert np.array_equal(getattr(got, f), getattr(hx, f)), f


def test_create_index_errors(tmp_path):
    hx = golden_index("tiny4k_5x4")
    sites = tmp_path / "sites.txt"
    sites.write_text(sites_text(hx))
    sc = tmp_path / "slices.txt"
    sc.write_text(masks_text(hx.masks))
    out = str(tmp_path / "o.issl")
    r = run([CREATE, str(sites), str(sc), "20"])
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = run([CREATE, str(sites), str(sc), "22", out])
    assert r.returncode == 1 and "greater than 21" in r.stderr
    r = run([CREATE, str(tmp_path / "missing.txt"), str(sc), "20", out])
    assert r.returncode == 1 and "Could not find" in r.stderr
    ragged = tmp_path / "ragged.txt"
    ragged.write_text(sites_text(hx) + "ACGT\n")
    r = run([CREATE, str(ragged), str(sc), "20", out])
    assert r.returncode == 1 and "not a multiple" in r.stderr
    assert not os.path.exists(out)


def test_create_index_21_positions_is_written_and_refused_by_the_scorer(tmp_path):
    """seqLength 21 is the builder's limit (ISSLCreateIndex.cpp:296); the
    scorer refuses such an index (its distance is taken on 20 positions)."""
    rng = np.random.default_rng(5)
    seqs = sorted({"".join("ACGT"[x] for x in rng.integers(0, 4, 21)) for _ in range(500)})
    sites = tmp_path / "sites.txt"
    sites.write_text("".join(s + "\n" for s in seqs))
    sc = tmp_path / "slices.txt"
    sc.write_text(masks_text([0xF << (4 * i) for i in range(5)], 21))
    out = str(tmp_path / "o.issl")
    r = run([CREATE, str(sites), str(sc), "21", out])
    assert r.returncode == 0, r.stderr
    got = issl.read_index(out)
    assert got.seq_len == 21 and got.n_sites == len(seqs)
    with pytest.raises(issl.IsslError) as e:
        issl.Index.open(out)
    assert e.value.code == issl._capi.ISSL_ERR_FORMAT
