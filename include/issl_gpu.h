/*
 * issl_gpu.h -- C ABI of the MI355X ISSL off-target scorer (libissl_gpu.so).
 *
 * The reference has no C ABI: its scorer lives inside
 * ISSLScoringModule::run (lib/ISSLScoringModule.cpp:52-466) and the
 * standalone ISSLScoreOfftargets main (ISSLScoreOfftargets/ISSLScoreOfftargets.cpp:36-647).
 * Each entry point below names the part of those it replaces.
 *
 * Conventions
 *   - Every function returns an int status: ISSL_OK (0) or a negative
 *     ISSL_ERR_* code; issl_last_error() then holds a message for the calling
 *     thread.  Nothing throws across the ABI.
 *   - Plain pointers and sizes only; buffers are caller-owned host memory
 *     unless a name says "_device".
 *   - Signatures are 2-bit packed 20-mers, position j in bits 2j..2j+1,
 *     A=0 C=1 G=2 T=3 (lib/ISSLScoringModule.cpp:10,469-476).
 *   - Method numbering is the pipeline enum of include/util.hpp:20.
 */
#ifndef ISSL_GPU_H
#define ISSL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define ISSL_OK              0
#define ISSL_ERR_ARG        -1  /* invalid argument                        */
#define ISSL_ERR_IO         -2  /* file could not be opened / read / written */
#define ISSL_ERR_FORMAT     -3  /* index file or arrays are inconsistent     */
#define ISSL_ERR_DEVICE     -4  /* HIP error, or no usable GPU               */
#define ISSL_ERR_NOMEM      -5  /* device or host allocation failed          */
#define ISSL_ERR_CAPACITY   -6  /* output capacity or admission limit exceeded */
#define ISSL_ERR_TIMEOUT    -7  /* a wait timed out                           */
#define ISSL_ERR_STATE      -8  /* call not valid in the object's state        */

/* --------------------------------------------------------------- methods */
/* include/util.hpp:20 */
#define ISSL_METHOD_MIT          0
#define ISSL_METHOD_CFD          1
#define ISSL_METHOD_MIT_AND_CFD  2
#define ISSL_METHOD_MIT_OR_CFD   3
#define ISSL_METHOD_AVG_MIT_CFD  4

/* Scoring forms.
 *   PLAIN   : one guide per workgroup, FP64, hits accumulated in the
 *             reference's slice-major posting order (bit-exact sums and the
 *             reference's early-exit point).
 *   BATCHED : persistent waves over (guide, chunk) tiles of the guide's
 *             slice-major posting sequence, sites streamed from a
 *             signature-inline device layout; precision selectable.
 *   ORDERED : the BATCHED launch structure with one tile, and so one wave,
 *             per guide: FP64 terms added one hit at a time in the
 *             reference's order with the exit test after each, so results
 *             are PLAIN's bit for bit (FP64 only).  A streamed index, or one
 *             whose guides could visit 2^26 - 64 postings or more, is scored
 *             by the PLAIN kernel instead (the same bits). */
#define ISSL_FORM_PLAIN    0
#define ISSL_FORM_BATCHED  1
#define ISSL_FORM_ORDERED  2

/* Precision of the BATCHED form (PLAIN is always FP64).
 *   FP64  : FP64 per-site terms and sums.
 *   FP32C : FP32 per-site products, each lane's sums as compensated
 *           two-float FP32 pairs; lane and tile partials, the transform and
 *           the threshold tests in FP64. */
#define ISSL_PREC_FP64   0
#define ISSL_PREC_FP32C  1

/* Flags */
#define ISSL_FLAG_DISTINCT  1u  /* also set the index's distinct-site bitmap
                                   (ISSLScoreOfftargets.cpp:202-208,375-377) */
#define ISSL_FLAG_GRAPH     2u  /* issl_score_device only: the first call runs
                                   the launch sequence and captures it in a
                                   hipGraph; later calls with the same buffers,
                                   stream, parameters and tuning replay it (no
                                   reference counterpart) */

typedef struct issl_index issl_index;

typedef struct {
    int32_t  max_dist;   /* must be 4: the reference tests dist <= 4
                            (ISSLScoringModule.cpp:292, ISSLScoreOfftargets.cpp:369) */
    int32_t  method;     /* ISSL_METHOD_* (util.hpp:20) */
    double   threshold;  /* score threshold; 0 disables the early exit
                            (maximum_sum = (10000 - 100 t)/t, ISSLScoringModule.cpp:225) */
    int32_t  form;       /* ISSL_FORM_* */
    int32_t  precision;  /* ISSL_PREC_* (BATCHED only) */
    uint32_t flags;      /* ISSL_FLAG_* */
    int32_t  chunk;      /* BATCHED tile length in postings; 0 = tuned/default */
} issl_params;

/* Per-guide results (each pointer nullable, n entries each). */
typedef struct {
    double   *mit;        /* 10000/(100+S_mit), -1 if MIT is not computed
                             (ISSLScoringModule.cpp:404-455) */
    double   *cfd;        /* 10000/(100+S_cfd), -1 if CFD is not computed */
    uint64_t *occ_count;  /* sum of occurrences of scored sites up to the exit
                             (ISSLScoreOfftargets.cpp:371, "Offtarget_Count") */
    uint64_t *site_count; /* number of distinct sites scored up to the exit */
    char     *pass;       /* '1' accepted, '0' rejected (lib/util.cpp:21-22) */
    double   *mit_sum;    /* raw S_mit before the transform */
    double   *cfd_sum;    /* raw S_cfd before the transform */
} issl_results;

typedef struct {
    uint64_t n_sites;      /* distinct sites N (header word 0) */
    uint64_t seq_len;      /* header word 1 */
    uint64_t n_slices;     /* header word 2 */
    uint64_t n_lists;      /* sum over slices of 4^w */
    uint64_t device_bytes; /* HBM held by the resident index */
    int32_t  device;       /* HIP device ordinal */
    int32_t  max_width;    /* largest slice width w */
} issl_index_info;

/* ------------------------------------------------------------------ index */

/* Load a .issl file and keep it resident on `device`.
 * Replaces the fread loader of ISSLScoringModule.cpp:72-150 and the list
 * pointer table of :181-198 (ISSLScoreOfftargets.cpp:103-243). */
int issl_index_open(const char *path, int device, issl_index **out);

/* Same, from arrays already in host memory in the .issl layout:
 * sigs[n_sites], slice_masks[n_slices], list_sizes[sum 4^w],
 * entries[n_slices*n_sites] (occ<<32 | id). */
int issl_index_from_arrays(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                           const uint64_t *sigs, const uint64_t *slice_masks,
                           const uint64_t *list_sizes, const uint64_t *entries,
                           int device, issl_index **out);

/* Build the slice lists on the device from distinct sites (id = array
 * position): for every slice a stable radix sort of the ids by slice key, the
 * list sizes and starts, and the posting entries occ<<32 | id, ids ascending
 * inside each list -- ISSLCreateIndex.cpp:450-484 done in HBM.  Only the
 * signatures (8 B) and occurrence counts (4 B) per site cross PCIe.  The
 * result is the same resident index issl_index_from_arrays makes from the
 * host-built lists. */
int issl_index_build(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                     const uint64_t *sigs, const uint32_t *occ,
                     const uint64_t *slice_masks, int device, issl_index **out);

/* Streamed index: the posting lists stay in host memory and pass through
 * slice buffers in HBM, so an index larger than the device memory can be
 * scored.  Replaces the memory-mapped loader of the reference's MMF scorer
 * (ISSLScoreOfftargetsMMF/ISSLScoreOfftargetsMMF.cpp:99-153,
 * lib/ISSLScoringModuleMMF.cpp:53-150): the file is memory-mapped and read
 * from the mapping.  Only the signatures (8 B per site), the list starts
 * and `slot_bytes` of slice buffers (16 B per site per buffer, at most
 * n_slices buffers; 0 takes as many as fit in the free device memory less
 * 4 GiB) stay on the device.  issl_score and issl_score_device stream the
 * slices through the buffers in ascending order, a window at a time, with
 * each guide's sums carried between windows: PLAIN results are bit-identical
 * to the resident index's, BATCHED within its tolerance.  With n_slices
 * buffers the slices stay resident after their first use.  Posting ids are
 * checked against n_sites while they are copied (ISSL_ERR_FORMAT).
 * Graph replay, candidate lists, validation, export and autotuning need a
 * resident index and return ISSL_ERR_ARG. */
int issl_index_open_streamed(const char *path, int device, uint64_t slot_bytes,
                             issl_index **out);

/* Same, over arrays in host memory in the .issl layout (as for
 * issl_index_from_arrays); `entries` must stay valid and unchanged until
 * the index is closed. */
int issl_index_from_arrays_streamed(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                                    const uint64_t *sigs, const uint64_t *slice_masks,
                                    const uint64_t *list_sizes, const uint64_t *entries,
                                    int device, uint64_t slot_bytes, issl_index **out);

/* Slice buffers of a streamed index, bytes copied into them so far and
 * windows scored (all 0 for a resident index).  No counterpart in the
 * reference. */
int issl_index_stream_stats(issl_index *ix, uint64_t *slots, uint64_t *uploaded_bytes,
                            uint64_t *windows);

/* Copy a resident index's lists back to the host in the .issl layout:
 * list_sizes[n_lists], entries[n_slices * n_sites] (for issl_write_index);
 * entries may be NULL to fetch the sizes only. */
int issl_index_export(issl_index *ix, uint64_t *list_sizes, uint64_t *entries);

/* Release the resident index (no counterpart: the reference frees its
 * function-local vectors at the end of run). */
int issl_index_close(issl_index *ix);

int issl_index_get_info(const issl_index *ix, issl_index_info *out);

/* Site-word layout of the BATCHED tile kernel.  No reference counterpart (the
 * reference reads a posting and then gathers the site's signature,
 * lib/ISSLScoringModule.cpp:241-292).  Every site in the list that a guide's
 * slice key selects holds the guide's bases at that slice's positions, so
 * only the other 20 - w positions can differ: the list-residual layout keeps
 * those positions' codes and a clipped occurrence count in 4 bytes per
 * posting (the signature-inline layout takes 8).  It is built when every
 * slice has at least 8 positions, all below 20.  A new index streams the
 * signature-inline words; issl_autotune times both layouts and keeps the
 * list-residual words unless the signature-inline words are faster by more
 * than 3% (the 4-byte words halve the bytes, which pays where the streams
 * are long enough to be bandwidth-bound); set overrides either.  Results are bit-identical either way for an index whose sites each
 * sit in the list of their own key (issl_index_validate); for one that does
 * not, only the signature-inline words (use = 0) take the distance on the
 * whole signature as the reference does.  get: *built, *in_use and
 * the occurrence bits a residual word holds (counts at or above 2^bits - 1
 * are read from the posting); each output may be NULL. */
int issl_index_set_residual(issl_index *ix, int32_t use);
int issl_index_get_residual(const issl_index *ix, int32_t *built, int32_t *in_use,
                            uint32_t *occ_bits);

/* Check on the device that every posting entry sits in the list of its own
 * key and that each slice lists every site once; *bad_entries receives the
 * number of violations.  (The reference trusts the file.) */
int issl_index_validate(issl_index *ix, uint64_t *bad_entries);

/* Distinct-site statistic of ISSLScoreOfftargets.cpp:549-555: sites scored
 * by any guide since the last reset (calls made with ISSL_FLAG_DISTINCT). */
int issl_distinct_count(issl_index *ix, uint64_t *out);
int issl_distinct_reset(issl_index *ix);

/* ---------------------------------------------------------------- scoring */

/* 20-character guides -> signatures.  text holds n guides, guide i at
 * text + i*stride (stride >= 20).  ISSLScoringModule.cpp:469-476,
 * ISSLScoreOfftargets.cpp:279-288.  Characters other than C, G, T encode as
 * A, as in the reference's table for upper-case input. */
int issl_encode(const char *text, uint64_t n, uint64_t stride, uint64_t *out);

/* Score n guides against a resident index; host in/out buffers, blocking.
 * Replaces the scoring loop of ISSLScoringModule.cpp:204-462 and
 * ISSLScoreOfftargets.cpp:296-505.  n == 0 is valid and does nothing. */
int issl_score(issl_index *ix, const uint64_t *guides, uint64_t n,
               const issl_params *params, issl_results *out);

/* Same with device buffers on a caller stream (hipStream_t passed as void*,
 * NULL = the index's own stream); asynchronous.  Each results pointer is a
 * device pointer (nullable).  Inputs must already be resident in HBM. */
int issl_score_device(issl_index *ix, const uint64_t *guides_device, uint64_t n,
                      const issl_params *params, issl_results *out_device,
                      void *stream);

/* Give a caller stream a scoring workspace of its own on this index, so that
 * issl_score_device calls on it overlap on the device with calls on the
 * index's stream and on other bound streams (calls on one stream stay in
 * order); unbound streams share the index's workspace, each call waiting for
 * the last one on another stream.  Overlapping batches fill the CUs that the
 * previous batch's planner, finalizer and last waves leave idle.  Each bound
 * stream holds its own workspace (tile plans, per-tile partials, queue: about
 * 95 MB for 10,000 guides, plus up to 0.8 GB of in-order split terms for the
 * ORDERED form) while bound; at most 8 per index; a resident index only.  Unbind waits for the
 * stream's calls on this index and frees the workspace: unbind a stream
 * before destroying it (closing the index waits for the whole device when a
 * bound stream is left).  Binding a bound stream again does nothing.  No
 * counterpart in the reference, which scores one guide list per process. */
int issl_index_bind_stream(issl_index *ix, void *stream);
int issl_index_unbind_stream(issl_index *ix, void *stream);

/* issl_score_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * BATCHED and ORDERED forms carve their scoring workspace (tile plans,
 * per-tile partials, the tile queue) from a 64 MiB device array that each
 * GPU the library runs on holds until the process ends; PLAIN needs none.
 * A call whose workspace would exceed it (about 27,000 guides at 10 slices;
 * the message gives the limit for the index) fails with ISSL_ERR_CAPACITY.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  A resident index only; ISSL_FLAG_GRAPH is refused.  Results equal
 * issl_score_device's bit for bit. */
int issl_score_device_noalloc(issl_index *ix, const uint64_t *guides_device, uint64_t n,
                              const issl_params *params, issl_results *out_device,
                              void *stream);

/* Candidate sets: for every guide the ids of the sites it scores with the
 * threshold disabled (dist <= 4, each site once, in the reference's
 * slice-major posting order), CSR form.  offsets[n+1]; ids[capacity].
 * Replaces the candidate stage of the reference's CUDA scorer
 * (ISSLScoreOfftargets/issl_cuda.cuh:46-63: distance scan + dedup by query).
 * Returns ISSL_ERR_CAPACITY (offsets still filled) if capacity is short. */
int issl_candidates(issl_index *ix, const uint64_t *guides, uint64_t n,
                    uint64_t *offsets, uint32_t *ids, uint64_t capacity);

/* ------------------------------------------------------------- autotuner */

typedef struct {
    int32_t chunk;        /* postings per tile (minimum; heavy guides get longer
                             tiles so no guide has more than 64) */
    int32_t block;        /* threads per workgroup (fixed at 256) */
    int32_t depth;        /* postings in flight per lane: 4, 8 or 16 */
    int32_t grid_per_cu;  /* persistent workgroups per CU: 1..32 */
    int32_t direct_depth; /* postings in flight per lane in the one-wave-per-
                             guide kernel: 4, 8 or 16 */
    int32_t heavy_kpost;  /* guides with more than heavy_kpost*1024 postings go
                             to the persistent tile kernel, the rest are scored
                             one wave per guide; 0 = all guides use tiles,
                             ISSL_HEAVY_NEVER = none do */
    double  us_per_guide; /* measured */
    /* ISSL_FORM_ORDERED's tile kernel (0 in any of the three: the default) */
    int32_t ord_depth;    /* postings in flight per lane: 4, 8 or 16 */
    int32_t ord_grid;     /* persistent workgroups per CU: 1..32 */
    int32_t ord_split;    /* with the exit off (threshold 0), guides with more
                             postings than this (>= 64) are split over tiles
                             of `chunk` postings that run in parallel and
                             write their hits' terms, which one wave then adds
                             in order; ISSL_SPLIT_NEVER = one tile per guide */
    int32_t ovl_grid;     /* persistent workgroups per CU of the tile kernels
                             (both forms) for launches on streams bound to
                             the index (issl_index_bind_stream), so that the
                             launches of several streams fit on the CUs at
                             once; 0 = 3.  Not timed by the autotuner (it
                             times one launch at a time) */
} issl_tuning;

#define ISSL_HEAVY_NEVER 0x7FFFFFFF
#define ISSL_SPLIT_NEVER 0x7FFFFFFF

/* Time the BATCHED kernel's launch shapes, and the site-word layouts the
 * index holds (issl_index_set_residual), on a sample of guides for this index
 * at `precision` (ISSL_PREC_*) and keep the fastest (also written to the
 * cache file when path is non-NULL; the cache is keyed by the index shape,
 * the sample size's power of two, the precision and the device).
 * issl_autotune tunes at ISSL_PREC_FP64.  No reference counterpart. */
int issl_autotune_precision(issl_index *ix, const uint64_t *sample_guides, uint64_t n,
                            int32_t precision, const char *cache_path, issl_tuning *best);
int issl_autotune(issl_index *ix, const uint64_t *sample_guides, uint64_t n,
                  const char *cache_path, issl_tuning *best);
int issl_get_tuning(issl_index *ix, issl_tuning *out);
int issl_set_tuning(issl_index *ix, const issl_tuning *in);

/* ------------------------------------------------------------ measurement */

/* Kernel timing (no reference counterpart; the reference prints wall-clock
 * phase timers, ISSLScoreOfftargets.cpp:185-188,296,507-510).  While enabled,
 * a pair of HIP events is attached to the dispatch of the dominant kernel of
 * every scoring launch on this index (k_plain for PLAIN, k_tiles or k_stream
 * for BATCHED, or k_stream and k_tiles together when long guides go to the
 * tile kernel; hipExtLaunchKernelGGL), so they time that kernel alone.
 * enable = N > 1 dispatches that kernel N times back to back (it is
 * idempotent; results are unchanged) between one recorded event pair, which
 * spreads the events' own overhead over N dispatches.  Enabling clears
 * earlier marks; read synchronizes and returns the summed kernel time and the
 * number of dispatches timed. */
int issl_timing_enable(issl_index *ix, int enable);
int issl_timing_read(issl_index *ix, double *kernel_ms, uint64_t *launches);

/* ------------------------------------------------------------- utilities */

/* Thread-local message for the last failing call on this thread. */
const char *issl_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int issl_device_count(void);

/* Library version string. */
const char *issl_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ISSL_GPU_H */
