"""Golden vectors for off-target extraction (ExtractOfftargets.cpp:109-356).

Writes two seeded input files and the output the reference would write for
them, computed with Python's regex engine running the reference's two
patterns verbatim (ExtractOfftargets.cpp:16-17), the reverse complement of
lib/util.cpp:50-63 and the input splitting of ExtractOfftargets.cpp:76-109:

  extract_genome.fa        multi-record FASTA: mixed case, N runs, CRLF and
                           blank lines, PAM-dense stretches, records shorter
                           than 23 bases
  extract_lines.txt        plain text, one sequence per line
  extract_expected.txt     sorted sites of both files, one per line,
                           duplicates kept

    python tests/golden/make_extract_golden.py
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FWD = re.compile(r"(?=([ACG][ACGT]{19}[ACGT][AG]G))")
BWD = re.compile(r"(?=(C[CT][ACGT][ACGT]{19}[TGC]))")
COMP = str.maketrans("ACGT", "TGCA")


def rc(s: str) -> str:
    return s[::-1].translate(COMP)


def sites_of(seq: str) -> list:
    out = [m.group(1)[:20] for m in FWD.finditer(seq)]
    out += [rc(m.group(1))[:20] for m in BWD.finditer(seq)]
    return out


def split(text: str) -> list:
    lines = text.split("\n")
    if text.endswith("\n"):
        lines = lines[:-1]
    ws = " \t\n\v\f\r"
    if lines and lines[0][:1] == ">":
        out, cur = [], []
        for ln in lines[1:]:
            ln = ln.strip(ws).upper()
            if ln[:1] == ">":
                out.append("".join(cur))
                cur = []
            else:
                cur.append(ln)
        out.append("".join(cur))
        return out
    return [ln.strip(ws).upper() for ln in lines] or [""]


def random_seq(rng, n, alphabet="ACGT", p=None):
    return "".join(rng.choice(list(alphabet), size=n, p=p))


def main():
    rng = np.random.default_rng(20241)
    recs = []
    recs.append((">chrA test record", random_seq(rng, 3000, "ACGT", [0.3, 0.2, 0.2, 0.3])))
    dense = random_seq(rng, 1500, "GCAT", [0.45, 0.35, 0.1, 0.1])
    recs.append((">chrB pam-dense", dense))
    mixed = list(random_seq(rng, 2000, "ACGTacgtN", [0.2, 0.2, 0.2, 0.2, 0.04, 0.04, 0.04, 0.04, 0.04]))
    recs.append((">chrC mixed case and N", "".join(mixed)))
    recs.append((">short", "ACGTACGTAGG"))
    recs.append((">empty", ""))
    # A repeat family: the same 23-mer many times (duplicates in the output).
    unit = "GACCTGATATACTAGGAACCAGG"
    recs.append((">chrD repeats", (unit + "TT") * 12 + random_seq(rng, 200)))
    lines = []
    for i, (hdr, seq) in enumerate(recs):
        lines.append(hdr)
        width = 60 if i % 2 == 0 else 71
        for k in range(0, len(seq), width):
            piece = seq[k:k + width]
            if i == 2 and k % 3 == 0:
                piece = "  " + piece + " \r"   # trimmed by the reference
            lines.append(piece)
        if i == 1:
            lines.append("")                   # a blank line inside a record
    fa = "\n".join(lines) + "\n"
    txt = "\n".join([random_seq(rng, 400), "acgt" * 30 + "ccaggatcgatcgatcgatcgatcga",
                     "", random_seq(rng, 22), random_seq(rng, 300, "ACGTN", [.24, .24, .24, .24, .04])]) + "\n"
    with open(os.path.join(HERE, "extract_genome.fa"), "w", newline="") as f:
        f.write(fa)
    with open(os.path.join(HERE, "extract_lines.txt"), "w", newline="") as f:
        f.write(txt)
    sites = []
    for text in (fa, txt):
        for seq in split(text):
            sites.extend(sites_of(seq))
    sites.sort()
    with open(os.path.join(HERE, "extract_expected.txt"), "w", newline="") as f:
        f.write("".join(s + "\n" for s in sites))
    print(f"{len(sites)} sites, {len(set(sites))} distinct")


if __name__ == "__main__":
    main()
