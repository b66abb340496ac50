/*
 * issl_index.h -- host-side .issl index construction and I/O (libissl_gpu.so).
 *
 * These replace the index builder ISSLCreateIndex/ISSLCreateIndex.cpp:284-493
 * (text parsing, occurrence counting, slice lists, file layout) and give the
 * scorer's loader a file reader.  Layout (all words little-endian u64):
 *   N | seqLength | S | sig[N] | sliceMask[S] |
 *   for each slice: size[4^w] then entry[N] = occ<<32 | id, grouped by key
 *   ascending and ids ascending within a key.   (ISSLCreateIndex.cpp:428-484)
 * Status codes and issl_last_error() as in issl_gpu.h.
 */
#ifndef ISSL_INDEX_H
#define ISSL_INDEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of lists of a slice configuration: sum over slices of 4^w, where w
 * counts the mask bits below seq_len (ISSLScoringModule.cpp:117-122).
 * Returns 0 when a slice is wider than 16 positions. */
uint64_t issl_list_count(const uint64_t *slice_masks, uint64_t n_slices,
                         uint64_t seq_len);

/* Slice lists from distinct sites (id = array position).
 * ISSLCreateIndex.cpp:450-484: for each slice, bucket ids by their slice key,
 * ids ascending inside a bucket, sizes[4^w] and entries[N] per slice.
 * list_sizes[issl_list_count()], entries[n_slices * n_sites].  threads <= 0
 * uses the OpenMP default. */
int issl_build_lists(const uint64_t *sigs, const uint32_t *occ, uint64_t n_sites,
                     uint64_t seq_len, const uint64_t *slice_masks,
                     uint64_t n_slices, uint64_t *list_sizes, uint64_t *entries,
                     int threads);

/* Write an index file (ISSLCreateIndex.cpp:428-484). */
int issl_write_index(const char *path, uint64_t n_sites, uint64_t seq_len,
                     uint64_t n_slices, const uint64_t *sigs,
                     const uint64_t *slice_masks, const uint64_t *list_sizes,
                     const uint64_t *entries);

/* Read the header and the list count of an index file (ISSLScoringModule.cpp:83-122). */
int issl_read_header(const char *path, uint64_t *n_sites, uint64_t *seq_len,
                     uint64_t *n_slices, uint64_t *n_lists);

/* Read a whole index file into caller buffers sized from issl_read_header
 * (ISSLScoringModule.cpp:93-147).  Fails with ISSL_ERR_FORMAT when the file
 * size does not match the header. */
int issl_read_index(const char *path, uint64_t *sigs, uint64_t *slice_masks,
                    uint64_t *list_sizes, uint64_t *entries);

/* Slice-mask text: one line per slice, '1' at position j puts j in the
 * slice (ISSLCreateIndex.cpp:345-365).  Writes up to cap masks; *n_out gets
 * the number of lines. */
int issl_parse_masks(const char *text, uint64_t len, uint64_t seq_len,
                     uint64_t *masks, uint64_t cap, uint64_t *n_out);

/* Sorted site text (one seq_len-mer per line, equal lines adjacent) ->
 * distinct signatures and occurrence counts, ids in file order
 * (ISSLCreateIndex.cpp:396-416).  line_len = seq_len + line-ending length.
 * Writes up to cap sites; *n_out gets the distinct count. */
int issl_count_sorted_sites(const char *text, uint64_t len, uint64_t seq_len,
                            uint64_t line_len, uint64_t *sigs, uint32_t *occ,
                            uint64_t cap, uint64_t *n_out);

/* Reference score tables as the product holds them, for checking:
 * mit_keys/mit_vals[6195] (fold masks ascending; MIT from the closed form of
 * ISSLCreateIndex.cpp:235-282), cfd_pos[320], cfd_pam[16]
 * (lib/otScorePenalties.cpp:6201-6546).  Each pointer nullable. */
int issl_score_tables(uint64_t *mit_keys, double *mit_vals, double *cfd_pos,
                      double *cfd_pam);

#ifdef __cplusplus
}
#endif
#endif /* ISSL_INDEX_H */
