/*
 * issl_oracle.h -- CPU restatement of the reference ISSL off-target scorer.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product (libissl_gpu.so) never links,
 * loads or calls it.
 *
 * Every function restates the reference's behaviour from its source; each
 * cites the file:line it follows.  How each function is pinned:
 *   or_encode            -- hand vectors + legacy fixture signatures
 *   or_mit_local         -- all 6,195 values of the legacy fixture's embedded
 *                           MIT table and of the compiled reference table
 *   or_cfd_tables        -- byte-equal to the compiled reference tables
 *                           (oracle/_ref) and to a committed SHA-256
 *   or_score             -- compiled reference ISSLScoringModule::run
 *                           (oracle/_ref) on generated indexes, committed as
 *                           golden vectors; brute force candidate sets
 *   or_candidates_brute  -- closed form over every site (independent of the
 *                           slice lists)
 *   or_extract           -- Python's regex engine running the reference's
 *                           two patterns verbatim (lookahead, every
 *                           position) on random and PAM-dense sequences,
 *                           and a committed golden output
 * Nothing here is "parity unpinned".
 */
#ifndef ISSL_ORACLE_H
#define ISSL_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Method numbering of include/util.hpp:20 (pipeline enum). */
enum { OR_MIT = 0, OR_CFD = 1, OR_MIT_AND_CFD = 2, OR_MIT_OR_CFD = 3, OR_AVG = 4 };

/* 20 characters -> 2-bit signature, position j in bits 2j..2j+1
 * (lib/ISSLScoringModule.cpp:10,469-476).  Characters outside the
 * 85-entry table read as 'A' here (the reference reads out of bounds). */
uint64_t or_encode(const char *seq20);

/* MIT local score from mismatch positions (ascending), closed form of
 * ISSLCreateIndex/ISSLCreateIndex.cpp:235-261. */
double or_mit_local(const uint64_t *positions, uint64_t k);

/* MIT score keyed by a fold mask (bit 2j set = mismatch at position j),
 * ISSLCreateIndex.cpp:271-282.  0.0 for an empty mask. */
double or_mit_fold(uint64_t fold_mask);

/* All 6,195 (fold mask, MIT) pairs for 1..4 mismatches, keys ascending.
 * Returns the count. keys/vals may be NULL to query the count. */
size_t or_mit_table(uint64_t *keys, double *vals);

/* CFD position (320) and PAM (16) tables, lib/otScorePenalties.cpp:6201-6546. */
void or_cfd_tables(double *pos320, double *pam16);

/* Restatement of the scoring loop of lib/ISSLScoringModule.cpp:204-462 and
 * ISSLScoreOfftargets/ISSLScoreOfftargets.cpp:300-505.
 *
 * Index arrays are the .issl layout (ISSLCreateIndex.cpp:428-484):
 *   sigs[n_sites], slice_masks[n_slices], list_sizes[sum_s 4^w_s],
 *   entries[n_slices * n_sites] = occ<<32 | id.
 * Outputs (each nullable):
 *   mit/cfd   : 10000/(100+sum) for the computed score, -1 otherwise
 *               (module semantics, ISSLScoringModule.cpp:404-455)
 *   occ_count : sum of occurrences of scored sites up to the exit
 *               (ISSLScoreOfftargets.cpp:371)
 *   site_count: number of sites scored (dist <= 4) up to the exit
 *   pass      : '1' accepted, '0' rejected (util.cpp:21-22)
 *   mit_sum/cfd_sum : the raw sums before the transform
 *   distinct  : (n_sites/64+1)-word bitmap, bit id set for every scored
 *               site (ISSLScoreOfftargets.cpp:375-377)
 *   margin    : over every hit up to the exit, the smallest relative
 *               distance |operand - cut| / |cut| of the exit test's operands
 *               (the sums, or their mean for avgMitCfd) from the cut; a sum
 *               order that differs only by rounding can move the exit only
 *               for guides whose margin is within that rounding.  +inf with
 *               the threshold disabled and for threshold 100 (cut 0: sums
 *               of nonnegative terms are 0 or positive in any order).
 * threads <= 0 uses the OpenMP default.  Returns 0, or -1 on a bad method
 * or index. */
int or_score(const uint64_t *sigs, uint64_t n_sites, uint64_t seq_len,
             const uint64_t *slice_masks, uint64_t n_slices,
             const uint64_t *list_sizes, const uint64_t *entries,
             const uint64_t *guides, uint64_t n_guides,
             int method, double threshold, int threads,
             double *mit, double *cfd, uint64_t *occ_count,
             uint64_t *site_count, char *pass,
             double *mit_sum, double *cfd_sum, uint64_t *distinct,
             double *margin);

/* Brute-force candidate set of one guide over every site:
 *   { id : popcount(fold(q ^ sig[id])) <= 4  and  exists s with
 *          fold & posmask2(s) == 0 }
 * where posmask2(s) has bit 2j for each position j of slice s.  Ids are
 * written ascending to ids_out (capacity cap); returns the set size (which
 * may exceed cap). */
uint64_t or_candidates_brute(const uint64_t *sigs, uint64_t n_sites,
                             uint64_t seq_len, const uint64_t *slice_masks,
                             uint64_t n_slices, uint64_t guide,
                             uint32_t *ids_out, uint64_t cap);

/* Off-target sites of one sequence, ExtractOfftargets/ExtractOfftargets.cpp:
 * 109-160.  Every position where the forward pattern (?=([ACG][ACGT]{19}
 * [ACGT][AG]G)) matches gives the 20 characters there; then every position
 * where the reverse pattern (?=(C[CT][ACGT][ACGT]{19}[TGC])) matches gives
 * the first 20 characters of the 23-mer's reverse complement (lib/util.cpp:
 * 50-63).  Letters are taken upper-cased.  Sites go to out as 20 characters
 * each (capacity cap sites), forward ones first; returns the number found
 * (which may exceed cap). */
uint64_t or_extract(const char *seq, uint64_t len, char *out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
