"""The CPU oracle against the reference's golden vectors and closed forms (no GPU)."""
import hashlib
import itertools
import os
import re

import numpy as np
import pytest

import oracle_lib as orc
from conftest import GOLDEN, GOLDEN_CASES, GOLDEN_THRESHOLDS, cut_of, golden_index, load_golden, tkey


def test_encode_hand_vectors():
    # position j -> bits 2j..2j+1, A=0 C=1 G=2 T=3 (lib/ISSLScoringModule.cpp:10,469-476)
    assert orc.lib().or_encode(b"A" * 20) == 0
    assert orc.lib().or_encode(b"C" + b"A" * 19) == 1
    assert orc.lib().or_encode(b"A" * 19 + b"T") == 3 << 38
    assert orc.lib().or_encode(b"ACGT" * 5) == sum(v << (2 * j) for j, v in
                                                  enumerate([0, 1, 2, 3] * 5))
    # non-ACGT upper-case letters read as A
    assert orc.lib().or_encode(b"N" * 20) == 0


def test_fold_worked_example():
    # lib/ISSLScoringModule.cpp:260-285: AATTGCAT vs ATATCGAT -> 4 mismatches
    q = orc.lib().or_encode(b"AATTGCAT" + b"A" * 12)
    o = orc.lib().or_encode(b"ATATCGAT" + b"A" * 12)
    x = q ^ o
    fold = ((x >> 1) & 0x5555555555555555) | (x & 0x5555555555555555)
    assert bin(fold).count("1") == 4
    assert [j for j in range(20) if (fold >> (2 * j)) & 3] == [1, 2, 4, 5]


def _legacy_mit():
    raw = np.fromfile(os.path.join(GOLDEN, "ref_test_genome_offTargets_indexed.issl"),
                      dtype=np.uint64)
    ns = int(raw[5])
    tab = raw[6:6 + 2 * ns].reshape(-1, 2)
    return tab[:, 0].copy(), tab[:, 1].copy().view(np.float64)


def test_mit_closed_form_matches_reference_fixture():
    """All 6,195 MIT values of the reference's test index, bit for bit."""
    lk, lv = _legacy_mit()
    k, v = orc.mit_table()
    assert k.shape == (6195,)
    assert (k == lk).all()
    assert (v.view(np.uint64) == lv.view(np.uint64)).all()


def test_mit_special_values():
    # one mismatch at position j: 100 * (1 - w[j]); w = 0 at 0,1,3,4,7
    for j, want in [(0, 100.0), (1, 100.0), (2, 98.6), (5, 60.5)]:
        assert orc.lib().or_mit_fold(1 << (2 * j)) == pytest.approx(want, rel=1e-15)
    assert orc.lib().or_mit_fold(0) == 0.0
    # a plausible slip (1/k instead of 1/k^2) fails this: two adjacent PAM-distal
    # mismatches at 0,1: T1 = 1, d = 1, T2 = 1/((18/19)*4+1), T3 = 1/4
    want = 1.0 * (1.0 / ((19.0 - 1.0) / 19.0 * 4.0 + 1)) * (1.0 / 4) * 100
    assert orc.lib().or_mit_fold(0b101) == want


def test_cfd_tables_checksum():
    """The 336 CFD constants, as the reference's compiled table bytes."""
    p, a = orc.cfd_tables()
    digest = hashlib.sha256(p.tobytes() + a.tobytes()).hexdigest()
    assert digest == CFD_SHA256
    # structural facts of the Doench table: matching pairs (rc of the guide
    # base) are 1, NGG PAM is 1, other PAMs < 1
    for pos in range(20):
        for g in range(4):
            assert p[(pos << 4) | (g << 2) | (g ^ 3)] == 1.0
    assert a[0b1010] == 1.0 and (a[np.arange(16) != 10] < 1).all()


# SHA-256 of cfdPosPenalties[320] || cfdPamPenalties[16] as doubles
# (lib/otScorePenalties.cpp:6201-6546), taken from the compiled reference.
CFD_SHA256 = "5d37e48d643fc7921ac514217b57bf3035ade136181b51b18d68cde7af209b33"


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_reference_goldens(name):
    """Bit-exact FP64 scores and identical calls for every method, threshold on and off."""
    gd = load_golden(name)
    ix = golden_index(name)
    g = gd["guides"]
    for m, t in itertools.product(range(5), GOLDEN_THRESHOLDS):
        r = orc.score(ix, g, m, t)
        key = f"m{m}_t{tkey(t)}"
        assert (r["mit"].view(np.uint64) == gd[key + "_mit"].view(np.uint64)).all(), key
        assert (r["cfd"].view(np.uint64) == gd[key + "_cfd"].view(np.uint64)).all(), key
        assert (r["pass"] == gd[key + "_pass"]).all(), key


@pytest.mark.parametrize("name", ["legacy39_5x4", "tiny4k_5x4", "mid50k_10x10"])
def test_oracle_candidate_sets_match_brute_force(name):
    """Threshold disabled: sites scored per guide == brute force over all sites."""
    gd = load_golden(name)
    ix = golden_index(name)
    g = gd["guides"][:200]
    r = orc.score(ix, g, 2, 0.0, distinct=True)
    occ = ix.occurrences()
    union = set()
    for i, q in enumerate(g):
        ids = orc.candidates_brute(ix, int(q))
        assert r["site_count"][i] == len(ids)
        assert r["occ_count"][i] == int(occ[ids].astype(np.uint64).sum())
        union.update(ids.tolist())
    bm = r["distinct"]
    assert sum(bin(int(w)).count("1") for w in bm) == len(union)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_negative_threshold_exits_at_first_hit(name):
    """t = -5: maximum_sum = -2,100 < 0 <= any running sum, so the loop stops
    right after the first site within distance 4 (lib/ISSLScoringModule.cpp:
    225,365-401): one site counted, its terms alone in the sums."""
    gd = load_golden(name)
    ix = golden_index(name)
    g = gd["guides"]
    full = orc.score(ix, g, 2, 0.0)
    for m in range(5):
        r = orc.score(ix, g, m, -5.0)
        assert (r["site_count"] == np.minimum(full["site_count"], 1)).all(), m
        assert (r["pass"] == ord("1")).all()   # scores are >= 0 > -5


@pytest.mark.parametrize("name", ["tiny4k_5x4", "mid50k_10x10"])
def test_exit_margin(name):
    """The margin of or_score bounds the distance of the final sums from the
    cut, is infinite with the exit disabled, and is 0 only at the cut."""
    gd = load_golden(name)
    ix = golden_index(name)
    g = gd["guides"]
    r0 = orc.score(ix, g, 2, 0.0, margin=True)
    assert np.isinf(r0["margin"]).all()
    for m, t in itertools.product(range(5), (75.0, 99.99, 1.0)):
        r = orc.score(ix, g, m, t, margin=True)
        cut = cut_of(t)
        fm = np.abs(r["mit_sum"] - cut) / cut
        fc = np.abs(r["cfd_sum"] - cut) / cut
        fa = np.abs((r["mit_sum"] + r["cfd_sum"]) / 2 - cut) / cut
        last = {0: fm, 1: fc, 2: np.minimum(fm, fc), 3: np.minimum(fm, fc), 4: fa}[m]
        hit = r["site_count"] > 0
        assert (r["margin"][hit] <= last[hit]).all(), (m, t)
        assert (r["margin"][hit] > 0).all(), (m, t)
        assert np.isinf(r["margin"][~hit]).all()


def test_threshold_zero_disables_exit():
    """t = 0: maximum_sum = +inf, never exits; pass never rejects (score < 0 is false)."""
    gd = load_golden("tiny4k_5x4")
    ix = golden_index("tiny4k_5x4")
    r = orc.score(ix, gd["guides"], 2, 0.0)
    assert (r["pass"] == ord("1")).all()
    # the sums are full: equal to the brute-force occurrence-weighted CFD count floor
    assert (r["cfd_sum"] >= 0).all()


# ---------------------------------------------------------------- extraction

_FWD = re.compile(r"(?=([ACG][ACGT]{19}[ACGT][AG]G))")
_BWD = re.compile(r"(?=(C[CT][ACGT][ACGT]{19}[TGC]))")
_COMP = str.maketrans("ACGT", "TGCA")


def _regex_sites(seq: str) -> list:
    """The reference's patterns (ExtractOfftargets.cpp:16-17) run by Python's
    regex engine: a lookahead match at every position."""
    out = [m.group(1)[:20] for m in _FWD.finditer(seq)]
    out += [m.group(1)[::-1].translate(_COMP)[:20] for m in _BWD.finditer(seq)]
    return out


@pytest.mark.parametrize("seed", range(6))
def test_extract_oracle_matches_reference_patterns(seed):
    rng = np.random.default_rng(seed)
    for trial in range(40):
        n = int(rng.integers(0, 900))
        if trial % 3 == 0:     # PAM-dense: overlapping matches on both strands
            alpha, p = list("GCAT"), [0.4, 0.4, 0.1, 0.1]
        elif trial % 3 == 1:   # mixed case and ambiguity codes
            alpha, p = list("ACGTacgtNRY"), [0.15] * 4 + [0.08] * 4 + [0.04, 0.02, 0.02]
        else:
            alpha, p = list("ACGT"), None
        seq = "".join(rng.choice(alpha, size=n, p=p))
        got = orc.extract(seq.encode())
        assert got == _regex_sites(seq.upper()), (seed, trial)


def test_extract_oracle_edge_cases():
    assert orc.extract(b"") == []
    assert orc.extract(b"ACGTACGTACGTACGTACGTAG") == []           # 22 bases
    assert orc.extract(b"TCGTACGTACGTACGTACGTAGG") == []          # T at the first base
    assert orc.extract(b"ACGTACGTACGTACGTACGTATG") == []          # PAM xTG
    assert orc.extract(b"ACGTACGTACGTACGTACGTAAG") == ["ACGTACGTACGTACGTACGT"]   # NAG
    assert orc.extract(b"acgtacgtacgtacgtacgtagg") == ["ACGTACGTACGTACGTACGT"]
    assert orc.extract(b"ACGTACGTACNTACGTACGTAGG") == []          # N inside the window
    # 'CCN' then 19 bases then [TGC] -> reverse complement
    assert orc.extract(b"CCAGTACGTACGTACGTACGTAC") == ["GTACGTACGTACGTACGTAC"]
    assert orc.extract(b"CCAGTACGTACGTACGTACGTAA") == []          # last base A


def test_extract_split_rules():
    # FASTA: headers start sequences, lines are trimmed, upper-cased, joined.
    assert orc.split_input(b">a\nacg \r\n tTT\n\n>b x\nGG\n") == [b"ACGTTT", b"GG"]
    # plain text: one sequence per line, the first line included
    assert orc.split_input(b"acg\n TT \r\n") == [b"ACG", b"TT"]
    assert orc.split_input(b"") == [b""]


def test_extract_oracle_golden():
    g = os.path.join(GOLDEN, "extract_")
    texts = [open(g + "genome.fa", "rb").read(), open(g + "lines.txt", "rb").read()]
    want = open(g + "expected.txt", "rb").read()
    assert orc.extract_output(texts) == want
