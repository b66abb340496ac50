"""Write the golden vectors of tests/golden/ from the reference scorer itself.

Runs in the development container only: it drives oracle/_ref/issl_ref_harness
(the reference's lib/ISSLScoringModule.cpp + lib/otScorePenalties.cpp built
from /root/reference by `make -C oracle ref`) on indexes made by this
package, and stores the reference's per-guide outputs
(mitOfftargetscore, cfdOfftargetscore, passedOffTargetScore) for every
method and for thresholds 0 (early exit disabled), 75, 100 (cut 0: exit at
the first positive sum), 99.99 (cut just above 0), 1 (cut 9,900) and -5
(cut -2,100: exit at the first hit).

Also copies two data files of the reference's own tests
(tests/data/test_genome_offTargets.fa and the legacy-layout
tests/data/test_genome_offTargets_indexed.issl) as fixtures.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import cracklingplusplus_gpu_amd as issl  # noqa: E402
from cracklingplusplus_gpu_amd import workloads  # noqa: E402

REF = "/root/reference"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "issl_ref_harness")
THRESHOLDS = (0.0, 75.0, 100.0, 99.99, 1.0, -5.0)


def tkey(t: float) -> str:
    """Key of a threshold in the golden files: 0, 75, 100, 99.99, 1, -5."""
    return str(int(t)) if t == int(t) else repr(t)
METHODS = (0, 1, 2, 3, 4)


def sha256_file(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def decode_legacy(path: str) -> issl.IndexArrays:
    """Legacy layout of the reference's test fixture:
    [N, seqLen, seqCount, sliceWidthBits, S, nScores], nScores (mask, f64)
    pairs, sig[N], all slices' sizes, all slices' entries.  Slices are the
    contiguous windows of sliceWidthBits/2 positions."""
    raw = np.fromfile(path, dtype=np.uint64)
    n, seq_len, _, wbits, s, ns = (int(x) for x in raw[:6])
    p = 6 + 2 * ns
    sigs = raw[p:p + n].copy()
    p += n
    w = wbits // 2
    sizes = raw[p:p + s * (1 << wbits)].copy()
    p += s * (1 << wbits)
    entries = raw[p:p + s * n].copy()
    masks = np.array([sum(1 << (w * i + j) for j in range(w)) for i in range(s)], dtype=np.uint64)
    return issl.IndexArrays(sigs, masks, sizes, entries, seq_len)


def run_ref(index_path: str, guides_txt: str, method: int, thr: float, n: int):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        subprocess.run([HARNESS, "score", index_path, guides_txt, str(method), repr(thr), "8", out],
                       check=True, stdout=subprocess.DEVNULL)
        raw = np.fromfile(out, dtype=np.uint8).reshape(n, 17)
    mit = raw[:, 0:8].copy().view(np.float64).ravel()
    cfd = raw[:, 8:16].copy().view(np.float64).ravel()
    passed = raw[:, 16].copy()
    return mit, cfd, passed


def golden_case(name: str, ix: issl.IndexArrays, guides: np.ndarray, index_path: str,
                kinds=None) -> None:
    seqs = issl.decode(guides)
    with tempfile.TemporaryDirectory() as td:
        gtxt = os.path.join(td, "guides.txt")
        with open(gtxt, "w") as f:
            for s in seqs:
                f.write(s + "NGG\n")
        data = {"guides": guides.astype(np.uint64),
                "issl_sha256": np.array(sha256_file(index_path)),
                "n_sites": np.uint64(ix.n_sites)}
        if kinds is not None:
            data["kinds"] = kinds.astype(np.uint8)
        for m in METHODS:
            for t in THRESHOLDS:
                mit, cfd, passed = run_ref(index_path, gtxt, m, t, len(guides))
                key = f"m{m}_t{tkey(t)}"
                data[key + "_mit"] = mit
                data[key + "_cfd"] = cfd
                data[key + "_pass"] = passed
    np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"), **data)
    print(name, len(guides), "guides")


def main() -> None:
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    # Fixtures held by the reference's own tests.
    for src in ("tests/data/test_genome_offTargets.fa",
                "tests/data/test_genome_offTargets_indexed.issl"):
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, "ref_" + os.path.basename(src)))

    with tempfile.TemporaryDirectory() as td:
        # The reference fixture, converted to the current layout.
        legacy = decode_legacy(os.path.join(HERE, "ref_test_genome_offTargets_indexed.issl"))
        path = os.path.join(td, "legacy39.issl")
        issl.write_index(path, legacy)
        rng = np.random.default_rng(39)
        sites = legacy.sigs
        queries = issl.encode(["ATCGATCGATCGATCGATCG", "GCTAGCTAGCTAGCTAGCTA",
                               "TAGCTAGCTAGCTAGCTAGC", "CGATCGATCGATCGATCGAT"])
        variants = sites.copy()
        for i in range(len(variants)):
            k = rng.integers(1, 5)
            for j in rng.choice(20, size=k, replace=False):
                c = (int(variants[i]) >> (2 * int(j))) & 3
                nc = (c + int(rng.integers(1, 4))) & 3
                variants[i] = (int(variants[i]) & ~(3 << (2 * int(j)))) | (nc << (2 * int(j)))
        rand = rng.integers(0, 1 << 40, size=60, dtype=np.uint64)
        guides = np.concatenate([sites, queries, variants, rand]).astype(np.uint64)
        golden_case("legacy39_5x4", legacy, guides, path)

        # Committed small index.
        sigs, occ, masks, g, kinds = workloads.make("tiny4k_5x4")
        ix = issl.build_lists(sigs, occ, masks)
        path = os.path.join(HERE, "tiny4k_5x4.issl")
        issl.write_index(path, ix)
        golden_case("tiny4k_5x4", ix, g, path, kinds)

        # Regenerated indexes (the npz records the .issl SHA-256).
        for name in ("mid50k_10x10", "small200k_8x8", "uniform1m_5x4"):
            sigs, occ, masks, g, kinds = workloads.make(name)
            ix = issl.build_lists(sigs, occ, masks)
            path = os.path.join(td, name + ".issl")
            issl.write_index(path, ix)
            golden_case(name, ix, g, path, kinds)


if __name__ == "__main__":
    main()
