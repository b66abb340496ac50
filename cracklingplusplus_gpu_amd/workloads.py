"""Seeded workload definitions (issl_synth configs).

Named after what they contain.  The large ones follow the benchmark
configurations: a 1 M-site uniform index with 5 contiguous 4-nt slices and
1,000 guides; a 100 M-site 41 %-GC index with 2 % repeat-derived sites and
10 seeded 10-position slices (10,000 or 100,000 guides); a 300 M-site index
with 5 % repeat-derived sites (50,000 guides).

Conventions fixed here:
  * "1 + Geometric(p)" counts failures before the first success, so most
    sites occur once (p = 0.8 -> mean 1.25);
  * seeded 10x10 slice masks are drawn without the covering requirement
    (10 masks cannot cover all 4,845 4-subsets of 20 positions; the
    Schoenheim bound for such a covering is 30 masks);
  * planted guides are uniform 20-mers; one variant of each with 1-4
    substitutions is added to the sites;
  * genome100m_10x10_heavy is the labelled heavy-tail variant: the stated
    generator of the 100 M-site configuration gives repeat-neighbourhood
    guides only about 4x the median posting volume, this one 100-1000x.
"""
from __future__ import annotations

from . import synth

BASE = dict(gc=0.5, repeat_frac=0.0, n_repeat_seeds=0, repeat_sub_p=0.1, occ_p_unique=0.8,
            occ_p_repeat=0.5, contiguous_masks=1, guide_batches=0, n_guides_site=0, n_guides_unique=0,
            n_guides_planted=0, n_guides_repeat=0, n_guides_random=0, guide_sub_p=0.0)

WORKLOADS = {
    # 1 M uniform sites (+200 planted variants), 5 contiguous 4-nt slices.
    "uniform1m_5x4": dict(BASE, n_sites=1_000_200, n_slices=5, slice_width=4,
                          n_guides_site=400, n_guides_planted=200, n_guides_random=400,
                          seed=0xC0C0),
    # 100 M sites at 41 % GC, 2 % from 2,000 repeat seeds, 10 seeded 10-position slices.
    "genome100m_10x10": dict(BASE, n_sites=100_000_000, gc=0.41, repeat_frac=0.02,
                             n_repeat_seeds=2000, repeat_sub_p=0.1, occ_p_repeat=0.5,
                             n_slices=10, slice_width=10, contiguous_masks=0,
                             n_guides_site=5000, n_guides_repeat=2500, n_guides_random=2500,
                             seed=0xC1C1),
    # The same index, 100,000 guides: 30 % repeat neighbourhoods, 70 % unique sites.
    "genome100m_10x10_q100k": dict(BASE, n_sites=100_000_000, gc=0.41, repeat_frac=0.02,
                                   n_repeat_seeds=2000, repeat_sub_p=0.1, occ_p_repeat=0.5,
                                   n_slices=10, slice_width=10, contiguous_masks=0,
                                   n_guides_unique=70_000, n_guides_repeat=30_000,
                                   seed=0xC1C1),
    # A heavy tail on a 100 M-site index: 4 % of the sites from two large repeat
    # families (2 M members each, Binomial(20, 0.2) substitutions), and 10,000
    # guides of which 1 % sit close to a family's seed (Binomial(20, 0.05)):
    # their lists hold 100-1000x the median posting volume.
    "genome100m_10x10_heavy": dict(BASE, n_sites=100_000_000, gc=0.41, repeat_frac=0.04,
                                   n_repeat_seeds=2, repeat_sub_p=0.2, guide_sub_p=0.05,
                                   occ_p_repeat=0.5, n_slices=10, slice_width=10,
                                   contiguous_masks=0, n_guides_unique=9_900,
                                   n_guides_repeat=100, seed=0xC1C1),
    # 300 M sites, 5 % from 10,000 repeat seeds, the 10x10 slices (same mask seed).
    "genome300m_10x10": dict(BASE, n_sites=300_000_000, gc=0.41, repeat_frac=0.05,
                             n_repeat_seeds=10_000, repeat_sub_p=0.1, occ_p_repeat=0.5,
                             n_slices=10, slice_width=10, contiguous_masks=0,
                             n_guides_site=25_000, n_guides_repeat=12_500,
                             n_guides_random=12_500, seed=0xC1C1),
    # Small cases for parity tests.
    "tiny4k_5x4": dict(BASE, n_sites=4000, repeat_frac=0.2, n_repeat_seeds=10,
                       repeat_sub_p=0.1, n_slices=5, slice_width=4, n_guides_site=100,
                       n_guides_planted=50, n_guides_repeat=50, n_guides_random=100,
                       seed=0x7171),
    "mid50k_10x10": dict(BASE, n_sites=50_000, gc=0.41, repeat_frac=0.1, n_repeat_seeds=50,
                         repeat_sub_p=0.1, n_slices=10, slice_width=10, contiguous_masks=0,
                         n_guides_site=200, n_guides_planted=50, n_guides_repeat=150,
                         n_guides_random=100, seed=0x5050),
    "small200k_8x8": dict(BASE, n_sites=200_000, gc=0.41, repeat_frac=0.05,
                          n_repeat_seeds=100, repeat_sub_p=0.1, n_slices=8, slice_width=8,
                          contiguous_masks=0, n_guides_site=400, n_guides_planted=100,
                          n_guides_repeat=300, n_guides_random=200, seed=0x2008),
}


def make(name: str, threads: int = 0, **override):
    """(sigs, occ, masks, guides, kinds) of a named workload."""
    cfg = dict(WORKLOADS[name])
    cfg.update(override)
    return synth(cfg, threads=threads)
