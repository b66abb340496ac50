/*
 * issl_server.h -- request loop over resident ISSL indexes on one device.
 *
 * No counterpart in the reference: its pipeline and CLI score one batch per
 * process run and re-read the index each time (lib/ISSLScoringModule.cpp:
 * 72-150, ISSLScoreOfftargets.cpp:103-183).  This serves many requests
 * against indexes kept in HBM:
 *
 *   * admission by free memory: an index is admitted only if the device's
 *     free HBM covers its size plus the configured reserve;
 *   * continuous batching into slots: a launcher thread fills each free
 *     slot (a stream with its own buffers) with up to max_batch guides from
 *     the oldest queued request and every queued request for the same index
 *     and parameters, splitting requests larger than a slot; up to `slots`
 *     launches are in flight at once, so requests arriving while others run
 *     are copied in and queued on the device as slots free up, and a
 *     completion thread returns results as each launch finishes;
 *   * admission by memory also for batches: a slot grows only if free HBM
 *     covers the growth plus the reserve, else its requests fail with
 *     ISSL_ERR_NOMEM;
 *   * isolation: results of a request never depend on what it was merged
 *     with (each guide's result depends on the guide alone), and a failing
 *     launch fails only the requests in it;
 *   * latency: per request, from submit to results in the caller's buffers;
 *     p50/p99 over the requests completed since the last reset.
 *
 * The caller's result buffers (issl_results, host memory) must stay valid
 * until the request completes, i.e. until issl_server_wait or issl_server_poll
 * reports it done; guides are copied at submit.  A request completes only
 * when no launch still holds any of its guides, also when it fails (a failed
 * launch, index removal, server shutdown): launches holding its other guides
 * still write their results into its buffers when they finish.
 */
#ifndef ISSL_SERVER_H
#define ISSL_SERVER_H

#include "issl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct issl_server issl_server;

typedef struct {
    int32_t  device;       /* HIP device ordinal */
    int32_t  slots;        /* launches in flight at once (own stream and buffers each);
                              0 = 3, at most 16 */
    uint64_t max_batch;    /* guides per launch; larger requests are split; 0 = 65536 */
    uint64_t hbm_reserve;  /* bytes of HBM left free when admitting an index or growing
                              a slot */
} issl_server_config;

typedef struct {
    uint64_t requests_done;    /* completed (ok or failed) since the last reset */
    uint64_t requests_failed;
    uint64_t guides_done;
    uint64_t launches;         /* scoring launches (merged batches) */
    uint64_t queued;           /* requests with guides not yet launched, now */
    uint64_t peak_inflight;    /* most launches in flight at once */
    uint64_t resident_bytes;   /* HBM held by the admitted indexes */
    uint64_t hbm_free, hbm_total;
    double   p50_ms, p99_ms, max_ms, mean_ms;   /* submit -> results in host buffers */
    double   mean_batch_guides;
} issl_server_stats;

int issl_server_create(const issl_server_config *cfg, issl_server **out);

/* Fails every request still queued (ISSL_ERR_STATE), waits for the launches
 * in flight (completing the requests they hold), closes the indexes and frees
 * the server. */
int issl_server_destroy(issl_server *srv);

/* Admit an index from a .issl file or from host arrays (the layout of
 * issl_index_from_arrays).  ISSL_ERR_NOMEM with an "admission" message when
 * free HBM does not cover the index plus the reserve. */
int issl_server_add_index(issl_server *srv, const char *path, int32_t *index_id);
int issl_server_add_index_arrays(issl_server *srv, uint64_t n_sites, uint64_t seq_len,
                                 uint64_t n_slices, const uint64_t *sigs,
                                 const uint64_t *slice_masks, const uint64_t *list_sizes,
                                 const uint64_t *entries, int32_t *index_id);

/* Remove an index: queued requests for it fail with ISSL_ERR_STATE (each
 * completes once no launch holds any of its guides); the launches in flight
 * on it complete first. */
int issl_server_remove_index(issl_server *srv, int32_t index_id);

/* Queue a request.  Parameters are checked here, so a bad request fails at
 * submit and never joins a batch.  *ticket identifies it for wait/poll. */
int issl_server_submit(issl_server *srv, int32_t index_id, const uint64_t *guides, uint64_t n,
                       const issl_params *params, const issl_results *out, uint64_t *ticket);

/* Block until the request completes (timeout_s < 0: no limit) and release its
 * ticket.  *status is the request's outcome (ISSL_OK or an error code; its
 * message is available from issl_server_wait's caller thread via
 * issl_last_error when non-zero).  Returns ISSL_ERR_STATE on timeout (the
 * ticket stays valid) or for an unknown ticket. */
int issl_server_wait(issl_server *srv, uint64_t ticket, double timeout_s, int32_t *status);

/* Non-blocking: *done = 1 and *status set when complete (ticket released). */
int issl_server_poll(issl_server *srv, uint64_t ticket, int32_t *done, int32_t *status);

/* Counters and latency percentiles; reset != 0 clears them afterwards. */
int issl_server_get_stats(issl_server *srv, int reset, issl_server_stats *out);

/* Free and total HBM of a device (hipMemGetInfo). */
int issl_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* HBM an index of this shape holds once resident (scoring workspace aside):
 * exact when every slice has 8 or more positions, else an upper bound (the
 * 4-byte list-residual words are then not built). */
uint64_t issl_index_bytes_needed(uint64_t n_sites, uint64_t n_slices, uint64_t n_lists);

#ifdef __cplusplus
}
#endif

#endif /* ISSL_SERVER_H */
