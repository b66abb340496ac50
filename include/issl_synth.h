/*
 * issl_synth.h -- seeded synthetic workloads (libissl_gpu.so).
 *
 * No reference counterpart: the reference obtains sites from a genome
 * (ExtractOfftargets/ExtractOfftargets.cpp:16-17,127-157 -> sorted text ->
 * ISSLCreateIndex).  These generators produce what that pipeline would hand
 * the scorer -- distinct 20-mer sites with ids in sorted-text order and
 * occurrence counts -- plus slice masks and query guides, so that benchmark
 * and test indexes can be made on any machine from a seed.
 *
 * Every output is a pure function of the config (the thread count does not
 * change it).  Geometric counts use the failures convention:
 * occ = 1 + F, P(F = k) = (1-p)^k p, so p = 0.8 gives mean 1.25.
 */
#ifndef ISSL_SYNTH_H
#define ISSL_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t n_sites;          /* distinct sites in the index, exactly */
    double   gc;               /* GC fraction of i.i.d. sites and random guides */
    double   repeat_frac;      /* fraction of n_sites drawn from repeat families */
    uint64_t n_repeat_seeds;   /* repeat seed 20-mers */
    double   repeat_sub_p;     /* per-position substitution prob. (Binomial(20,p)) */
    double   occ_p_unique;     /* occurrences of i.i.d. sites: 1 + Geometric(p) */
    double   occ_p_repeat;     /* occurrences of repeat-derived sites */
    uint64_t n_slices;         /* slice masks */
    uint64_t slice_width;      /* positions per mask */
    int32_t  contiguous_masks; /* 1: windows 0..w-1, w..2w-1, ...; 0: seeded draws */
    int32_t  guide_batches;    /* disjoint guide batches of the stated mix (0 = 1): batch b
                                  draws items b*count .. b*count + count - 1 of each guide
                                  stream and is shuffled on its own; batch 0 is the single
                                  batch.  Planted guides are the same in every batch (their
                                  variants are sites). */
    uint64_t n_guides_site;    /* guides copied from the site set */
    uint64_t n_guides_unique;  /* guides copied from non-repeat sites */
    uint64_t n_guides_planted; /* random guides; a 1-4 substitution variant of
                                  each is added to the sites */
    uint64_t n_guides_repeat;  /* repeat seed + Binomial(20, repeat_sub_p) subs */
    uint64_t n_guides_random;  /* i.i.d. at gc */
    uint64_t seed;
    double   guide_sub_p;      /* per-position substitution prob. of repeat-neighbourhood
                                  guides (0: repeat_sub_p, as for the family's sites) */
} issl_synth_cfg;

/* Generate one workload.  Buffers: sigs[n_sites] and occ[n_sites] (ids in
 * sorted-text order, all distinct), masks[n_slices],
 * guides[max(1, guide_batches) * sum of n_guides_*] (batch after batch, each
 * shuffled), kinds[same] (nullable; 0 site, 1 unique, 2 planted, 3 repeat,
 * 4 random). */
int issl_synth(const issl_synth_cfg *cfg, uint64_t *sigs, uint32_t *occ,
               uint64_t *masks, uint64_t *guides, uint8_t *kinds, int threads);

#ifdef __cplusplus
}
#endif
#endif /* ISSL_SYNTH_H */
