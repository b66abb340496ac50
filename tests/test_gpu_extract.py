"""Off-target extraction on the GPU (issl_extract.h) against the golden
extraction fixtures (made with Python's regex engine running the reference's
patterns) and the CPU oracle (MI355X).  Every comparison is exact: the
sorted site list, the distinct sites and their counts, the output file's
bytes, and the index built from the extracted sites."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cracklingplusplus_gpu_amd as issl
import oracle_lib as orc
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

FA = os.path.join(GOLDEN, "extract_genome.fa")
TXT = os.path.join(GOLDEN, "extract_lines.txt")
EXPECTED = os.path.join(GOLDEN, "extract_expected.txt")
CLI = os.path.join(ROOT, "cracklingplusplus_gpu_amd", "bin", "ExtractOfftargets")


def distinct_counts(lines):
    """(signatures, occurrences) of a sorted site list, as ISSLCreateIndex
    counts them."""
    text = "".join(x + "\n" for x in lines).encode()
    return issl.count_sorted_sites(text)


def random_genome(rng, n, dense=False):
    if dense:
        return bytes(rng.choice(list(b"GCATgcN"), size=n, p=[.35, .35, .1, .1, .04, .04, .02]))
    return bytes(rng.choice(list(b"ACGT"), size=n, p=[.3, .2, .2, .3]))


@pytest.mark.parametrize("chunk", [0, 64, 100, 1000, 4096])
def test_extract_golden_files(gpu_available, tmp_path, chunk):
    print(f"[gpu-extract] golden chunk {chunk}", flush=True)
    out = tmp_path / "out.txt"
    with issl.Extractor(chunk_bytes=chunk) as ex:
        ex.add_path(FA)
        ex.add_path(TXT)
        n, d = ex.finish()
        ex.write(str(out))
        sigs, occ = ex.sites()
    want = open(EXPECTED, "rb").read()
    assert out.read_bytes() == want
    assert n == want.count(b"\n")
    ws, wo = issl.count_sorted_sites(want)
    assert d == ws.shape[0]
    assert (sigs == ws).all() and (occ == wo).all()


@pytest.mark.parametrize("seed", range(4))
def test_extract_random_vs_oracle(gpu_available, seed):
    rng = np.random.default_rng(100 + seed)
    seqs = [random_genome(rng, int(rng.integers(0, 30000)), dense=bool(k % 2)) for k in range(12)]
    seqs += [b"", b"ACGT", random_genome(rng, 23), random_genome(rng, 200_000)]
    want = sorted(s for q in seqs for s in orc.extract(q))
    ws, wo = distinct_counts(want)
    for chunk in (0, 777, 8192):
        print(f"[gpu-extract] random seed {seed} chunk {chunk}", flush=True)
        with issl.Extractor(chunk_bytes=chunk) as ex:
            for q in seqs:
                ex.add_sequence(q)
            n, d = ex.finish()
            sigs, occ = ex.sites()
            bases, sites = ex.stats()
        assert n == len(want) and sites == len(want)
        assert bases == sum(len(q) for q in seqs)
        assert d == ws.shape[0]
        assert (sigs == ws).all() and (occ == wo).all()


def test_extract_large_genome_vs_oracle(gpu_available, tmp_path):
    """A 12 Mbase genome with a repeat family, scanned in 4 MiB pieces."""
    rng = np.random.default_rng(7)
    g = bytearray(random_genome(rng, 12_000_000))
    unit = random_genome(rng, 40)
    for p in rng.integers(0, len(g) - 40, size=3000):
        g[int(p):int(p) + 40] = unit
    g = bytes(g)
    sites = orc.extract(g)
    want_sorted = sorted(sites)
    ws, wo = distinct_counts(want_sorted)
    out = tmp_path / "big.txt"
    with issl.Extractor(chunk_bytes=4 << 20) as ex:
        ex.add_sequence(g)
        n, d = ex.finish()
        sigs, occ = ex.sites()
        ex.write(str(out))
    assert n == len(sites) and d == ws.shape[0]
    assert (sigs == ws).all() and (occ == wo).all()
    assert out.read_bytes() == "".join(x + "\n" for x in want_sorted).encode()


def test_extracted_index_matches_host_pipeline(gpu_available, tmp_path):
    """Sites -> index on the device equals the reference's route: the sorted
    file, ISSLCreateIndex's counting and slice lists (the host builder)."""
    rng = np.random.default_rng(3)
    masks = issl.parse_masks("\n".join(
        "".join("1" if j in range(5 * s, 5 * s + 5) else "0" for j in range(20)) for s in range(4)))
    out = tmp_path / "sites.txt"
    with issl.Extractor() as ex:
        ex.add_sequence(random_genome(rng, 2_000_000))
        ex.add_sequence(random_genome(rng, 500_000, dense=True))
        ex.finish()
        ex.write(str(out))
        with ex.build_index(masks) as dix:
            sizes, entries = dix.export()
    sigs, occ = issl.count_sorted_sites(out.read_bytes())
    hx = issl.build_lists(sigs, occ, masks)
    assert (sizes == hx.list_sizes).all()
    assert (entries == hx.entries).all()


def test_extract_edge_cases(gpu_available, tmp_path):
    with issl.Extractor() as ex:
        assert ex.finish() == (0, 0)
        sigs, occ = ex.sites()
        assert sigs.shape == (0,) and occ.shape == (0,)
        out = tmp_path / "empty.txt"
        ex.write(str(out))
        assert out.read_bytes() == b""
        ex.add_sequence(b"NNNNNNNNNNNNNNNNNNNNNNNNNNNNN")
        ex.add_sequence(b"acgtacgtacgtacgtacgtagg")        # lower case
        ex.add_sequence("ACGTACGTACGTACGTACGTAG")            # 22 bases
        assert ex.finish() == (1, 1)
        ex.add_sequence(b"ACGTACGTACGTACGTACGTAGG")          # adding after finish
        n, d = ex.finish()
        assert (n, d) == (2, 1)
        sigs, occ = ex.sites()
        assert occ.tolist() == [2]
        assert issl.decode(sigs) == ["ACGTACGTACGTACGTACGT"]
    with pytest.raises(issl.IsslError, match="finish"):
        with issl.Extractor() as ex:
            ex.sites()


def test_extract_directory_and_missing_path(gpu_available, tmp_path, capfd):
    d = tmp_path / "inputs"
    d.mkdir()
    shutil.copy(FA, d / "a.fa")
    shutil.copy(TXT, d / "b.txt")
    (d / "sub").mkdir()
    out = tmp_path / "out.txt"
    with issl.Extractor() as ex:
        ex.add_path(str(d))
        ex.add_path(str(tmp_path / "missing.fa"))
        ex.finish()
        ex.write(str(out))
    assert out.read_bytes() == open(EXPECTED, "rb").read()
    assert "Error processing:" in capfd.readouterr().out


def test_extract_cli_matches_golden(gpu_available, tmp_path):
    out = tmp_path / "offtargets.txt"
    r = subprocess.run([CLI, str(out), FA, TXT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == open(EXPECTED, "rb").read()
    assert "Done" in r.stdout
