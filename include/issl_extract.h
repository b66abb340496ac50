/*
 * issl_extract.h -- off-target site extraction on the GPU, the step before
 * the index: the reference's ExtractOfftargets
 * (ExtractOfftargets/ExtractOfftargets.cpp:19-356), which scans genome
 * sequences for sites followed by an NAG/NGG PAM on either strand and
 * writes all of them, sorted, to a text file that ISSLCreateIndex reads
 * (ISSLCreateIndex/ISSLCreateIndex.cpp:396-421).
 *
 * Sequences are scanned in HBM as they are added; finish sorts the sites in
 * text order and counts repeats on the device.  The result can be written
 * as the reference's file, read back as (signature, occurrences) pairs, or
 * turned into a resident index without leaving the device.
 *
 * Status codes and issl_last_error() are those of issl_gpu.h.  One
 * extractor is used by one thread at a time.
 */
#ifndef ISSL_EXTRACT_H
#define ISSL_EXTRACT_H

#include <stdint.h>

#include "issl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct issl_extractor issl_extractor;

/* New extractor on `device`.  chunk_bytes: bases scanned per launch (0:
 * 256 MiB); longer sequences are scanned in pieces overlapping by 22 bases
 * and shorter ones are packed together.  No counterpart in the reference
 * (its state is a temporary directory, ExtractOfftargets.cpp:29-31). */
int issl_extractor_create(int device, uint64_t chunk_bytes, issl_extractor **out);
int issl_extractor_destroy(issl_extractor *ex);

/* Scan one sequence (ExtractOfftargets.cpp:109-160): every position where
 * (?=([ACG][ACGT]{19}[ACGT][AG]G)) matches gives the 20 bases there, every
 * position where (?=(C[CT][ACGT][ACGT]{19}[TGC])) matches gives the first 20
 * bases of the 23-mer's reverse complement (lib/util.cpp:50-63).  Letters
 * count upper-cased; any other byte ends every window through it, so no
 * site spans two sequences.  The bases are copied; `seq` may be reused on
 * return. */
int issl_extractor_add_sequence(issl_extractor *ex, const char *seq, uint64_t len);

/* Read one input path as ExtractOfftargets.cpp:32-109 does and scan its
 * sequences: a regular file whose first line starts with '>' is FASTA (every
 * later line trimmed; a '>' line starts the next sequence, other lines are
 * appended); any other regular file holds one sequence per line (trimmed);
 * a directory adds the regular files directly in it.  A path that is
 * neither is skipped with the reference's message on stdout. */
int issl_extractor_add_path(issl_extractor *ex, const char *path);

/* Sort every site added so far in text order and count repeats
 * (ExtractOfftargets.cpp:161-356, the counting of ISSLCreateIndex.cpp:
 * 396-421).  Adding sequences afterwards and finishing again covers all of
 * them.  At most 2^31 - 1 sites per finish. */
int issl_extractor_finish(issl_extractor *ex, uint64_t *n_sites, uint64_t *n_distinct);

/* After finish: the distinct sites in text order, as signatures (position j
 * in bits 2j, lib/ISSLScoringModule.cpp:469-476) with their occurrence
 * counts -- the ids and occurrences ISSLCreateIndex assigns.  Either array
 * may be NULL. */
int issl_extractor_sites(issl_extractor *ex, uint64_t *sigs, uint32_t *occ);

/* After finish: write the reference's output file, every site as 20 bases
 * and '\n', sorted, repeats kept (ExtractOfftargets.cpp:290-340); the text
 * is formed on the device.  An extraction with no sites writes an empty
 * file (the reference fails there). */
int issl_extractor_write(issl_extractor *ex, const char *path);

/* After finish: a resident index over the distinct sites, built on the
 * device from the sites in HBM (the lists of issl_index_build).  seq_len is
 * 20. */
int issl_extractor_build_index(issl_extractor *ex, const uint64_t *slice_masks, uint64_t n_slices,
                               issl_index **out);

/* Bases scanned and sites found so far (no counterpart in the reference). */
int issl_extractor_stats(issl_extractor *ex, uint64_t *bases, uint64_t *sites);

#ifdef __cplusplus
}
#endif

#endif /* ISSL_EXTRACT_H */
