// device.hpp -- resident device index and kernel interfaces of libissl_gpu.so.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "issl_gpu.h"

namespace issl {

#define ISSL_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::issl::fail(ISSL_ERR_DEVICE, std::string(#call) + ": " +             \
                                                     hipGetErrorString(e_));             \
    } while (0)

struct StreamSource;  // host side of a streamed index (stream.hip)

// Per slice: fold mask (bit 2j per position j), first list-start index,
// width and positions.
struct SliceMeta {
    uint64_t fold;
    uint32_t lbase;
    uint32_t width;
    uint32_t keep;   // bit p set for each position p < 20 the slice does not fix
    uint8_t pos[16];
};

// Slice key of a signature: the 2-bit codes at the slice's positions,
// ascending, concatenated low to high (ISSLScoringModule.cpp:235-239).  The
// positions are read off the slice's fold mask (bit 2j for position j).
__device__ __forceinline__ uint32_t slice_key(uint64_t sig, uint64_t fold) {
    uint32_t key = 0;
    for (int j = 0; fold; ++j) {
        const int p = __builtin_ctzll(fold);
        key |= static_cast<uint32_t>((sig >> p) & 3ULL) << (2 * j);
        fold &= fold - 1;
    }
    return key;
}

// heavy_kpost value meaning "no guide goes to the tile kernel".
constexpr int32_t kHeavyNever = ISSL_HEAVY_NEVER;
// ORDERED: guides with more postings than this are split over tiles
// (issl_tuning.ord_split); term entries the workspace holds for them at most.
constexpr int32_t kSplitNever = ISSL_SPLIT_NEVER;
constexpr int32_t kOrdSplitDefault = 65536;
// Caller streams one index gives workspaces of their own (issl_index_bind_stream).
constexpr int kMaxBoundStreams = 8;
// Guides per piece of a host-buffer call (issl_score) on a resident index.
constexpr uint64_t kHostChunk = 1ULL << 18;
// Tile-kernel workgroups per CU of launches on bound streams (issl_tuning.ovl_grid 0).
constexpr int32_t kOvlGridDefault = 3;
constexpr uint64_t kOrdTermCap = 1ULL << 25;

// Site word of the signature-inline layout: signature in bits 0..39,
// occurrences in bits 40..63; kOccEscape means "read occ from the posting".
constexpr uint64_t kSigBits = (1ULL << 40) - 1;
constexpr uint32_t kOccEscape = 0xFFFFFF;

// List-residual site word (BATCHED tile kernel): every site in the list of
// slice s for key k holds k's bases at the slice's positions, which are the
// guide's own when the guide selected that list, so only the other 20 - w_s
// positions can differ.  Their 2-bit codes, ascending by position, fill bits
// 0 .. 2 r_s - 1 (r_s = 20 - w_s); bits res_shift .. 31 hold
// min(occ, escape), escape = 2^(32 - res_shift) - 1 meaning "read occ from
// the posting"; res_shift = 2 max_s r_s.  Built when every slice lies within
// positions 0..19 and has at least kResMinWidth positions (at least 8
// occurrence bits).
constexpr int kResMinWidth = 8;
constexpr int kResPositions = 20;
constexpr int kResMaxCodes = kResPositions - kResMinWidth;   // residual positions per slice

// The codes of `sig` at the positions 0..19 outside `fold` (bit 2j per slice
// position j), ascending, packed from bit 0: a list-residual word's code
// field, and a guide's residual for the lists it selects.
__host__ __device__ __forceinline__ uint32_t residual_code(uint64_t sig, uint64_t fold) {
    uint32_t r = 0, j = 0;
#pragma unroll
    for (int p = 0; p < kResPositions; ++p) {
        const uint32_t keep = static_cast<uint32_t>(~(fold >> (2 * p))) & 1u;
        r |= (static_cast<uint32_t>(sig >> (2 * p)) & 3u & (0u - keep)) << (2 * j);
        j += keep;
    }
    return r;
}

// residual_code from the slice's kept positions (SliceMeta::keep): one step
// per kept position instead of one per position; the same value.
__device__ __forceinline__ uint32_t residual_code_kept(uint64_t sig, uint32_t keep) {
    uint32_t r = 0, j = 0;
    for (uint32_t m = keep; m; m &= m - 1, j += 2) {
        const uint32_t p = static_cast<uint32_t>(__builtin_ctz(m));
        r |= (static_cast<uint32_t>(sig >> (2 * p)) & 3u) << j;
    }
    return r;
}

// Device CFD table: 320 position factors, 16 PAM factors, then 1.0.
constexpr int kCfdTable = 337;

// Grow-only device buffer; freed when it goes out of scope.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int reserve(size_t want);
    void release();
};

// Grow-only pinned host buffer; freed when it goes out of scope.
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { release(); }
    int reserve(size_t want);
    void release();
};

// Device outputs of one scoring call (each may be null).
struct DevResults {
    double *mit = nullptr, *cfd = nullptr, *mit_sum = nullptr, *cfd_sum = nullptr;
    uint64_t *occ = nullptr, *sites = nullptr;
    char *pass = nullptr;
};

// Scalar parameters shared by the kernels.
struct ScoreScalars {
    int method;
    int calc_mit;
    int calc_cfd;
    int use_cut;       // the early exit is armed (threshold != 0)
    double cut;        // (10000 - 100 t)/t, formed on the host in FP64
    double threshold;
};

struct Workspace {
    DevBuf guides, out, tiles, scratch;
    DevBuf state;      // carried guide state of a streamed index (two copies)
    HostBuf pinned;
    // The tile queue's counters in `tiles` may be nonzero: a launch sequence
    // stopped before its finalizer was queued.
    bool queue_dirty = false;
    // A caller's device buffer to carve the BATCHED / ORDERED workspace from
    // instead of `tiles` (nothing is allocated; a launch needing more than
    // ext_bytes fails with ISSL_ERR_CAPACITY; its queue counters are cleared
    // by a fill kernel at the start of every launch sequence).
    char *ext = nullptr;
    size_t ext_bytes = 0;
};

// Running state of one guide between slice windows: the two sums in FP64,
// the occurrence and site counts, and in bit 63 of `sites` whether the early
// exit has been taken.
struct GuideState {
    double m, c;
    uint64_t occ;
    uint64_t sites;
};
constexpr uint64_t kExited = 1ULL << 63;

// Slices [s0, s1) of an index whose postings sit at `post` / `inl`, slice s
// at (s - s0) * N.  A launch over a window reads the guide state from
// `state_in` (null: start from zero) and writes it to `state_out`, or, when
// `state_out` is null, the final results.  The resident index is the single
// window [0, S) over its own arrays with no state.
struct Window {
    int s0 = 0, s1 = 0;
    const uint64_t *post = nullptr;
    const uint64_t *inl = nullptr;
    const uint32_t *res = nullptr;   // list-residual words (resident index only), or null
    uint32_t res_shift = 0;
    const GuideState *state_in = nullptr;
    GuideState *state_out = nullptr;
};

// A launch sequence of issl_score_device captured in a hipGraph, with
// everything it bakes in (buffers, parameters, tuning, workspace address).
struct GraphEntry {
    const uint64_t *guides;
    uint64_t n;
    issl_params params;
    issl_results out;
    hipStream_t stream;
    issl_tuning tuning;
    const void *ws;
    hipGraphExec_t exec;
};

}  // namespace issl

struct issl_index {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0, seq_len = 0, S = 0, n_lists = 0;
    uint64_t p_max = 0;             // sum over slices of the longest list: bounds P(g)
    int max_w = 0;
    std::vector<uint64_t> masks;
    std::vector<issl::SliceMeta> meta_h;
    // device arrays
    uint64_t *d_sig = nullptr;      // [n] signatures
    uint64_t *d_post = nullptr;     // [S*n] occ<<32 | id, .issl order
    uint64_t *d_inline = nullptr;   // [S*n] sig | occ<<40 (same order)
    uint32_t *d_res = nullptr;      // [S*n] list-residual words (same order), or null
    uint32_t res_shift = 0;         // first occurrence bit of a residual word
    bool res_use = false;           // BATCHED streams d_res (when built): autotuner or caller
    uint32_t *d_lstart = nullptr;   // per slice: 4^w + 1 list starts
    issl::SliceMeta *d_meta = nullptr;
    double *d_mit64 = nullptr;      // [6195] by combinadic rank
    float *d_mit32 = nullptr;
    double *d_mitf = nullptr;       // [kMitFactorEntries] the closed form's factors
    double *d_cfd64 = nullptr;      // [320 pos | 16 pam | 1.0]
    float *d_cfd32 = nullptr;
    uint64_t *d_distinct = nullptr; // [n/64 + 1]
    uint64_t distinct_words = 0;
    uint64_t device_bytes = 0;
    issl_tuning tuning{};
    bool tuned = false;
    // Load depth and workgroups per CU of the ORDERED form's tile kernel
    // (one tile per guide): issl_autotune times them beside the BATCHED
    // shapes, issl_set_tuning sets them with the BATCHED ones.
    int32_t ord_depth = 4, ord_grid = 6, ord_split = issl::kOrdSplitDefault;
    // Tile-kernel workgroups per CU of launches on bound streams (0: the default).
    int32_t ovl_grid = 0;
    // Kernel timing: event pairs around the dominant kernel of each launch.
    int timing = 0;                 // 0 off; else dispatches per event pair
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::mutex mu;
    issl::Workspace ws;
    std::vector<issl::GraphEntry> graphs;   // captured sequences (ISSL_FLAG_GRAPH)
    // The scoring workspace is shared by every call on the index: a launch on
    // another stream than the previous one first waits for that one's work.
    // After a call on the index's own stream the event is recorded only when
    // a call on another stream comes (that stream outlives the index's calls;
    // a caller's stream is recorded at once, as it may be destroyed).
    bool ws_used = false;
    bool ws_lazy = false;
    hipStream_t ws_stream = nullptr;
    hipEvent_t ws_done = nullptr;
    // Caller streams bound with issl_index_bind_stream, each with a workspace
    // of its own: calls on a bound stream use it (the stream orders them) and
    // overlap on the device with calls on every other stream.
    struct BoundWs {
        hipStream_t stream = nullptr;
        issl::Workspace ws;
    };
    std::vector<std::unique_ptr<BoundWs>> bound;
    // Streamed index (issl_index_open_streamed and issl_index_from_arrays_streamed):
    // d_post and d_inline are null and the postings come from here.
    issl::StreamSource *src = nullptr;
};

namespace issl {

// Launchers (device.hip / batched.hip).  All asynchronous on `stream`.
// `win` null: the whole resident index.
int launch_plain(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                 const DevResults &out, uint64_t *distinct, const uint64_t *cand_off,
                 uint32_t *cand_ids, hipStream_t stream, const Window *win = nullptr);
// Bytes of the BATCHED / ORDERED workspace one launch sequence carves for n
// guides over S slices (res: with the list-residual codes) in a caller's
// workspace (Workspace::ext), which holds no in-order split terms.
size_t batch_ws_bytes(uint64_t n, int S, bool res);
int launch_batched(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                   int precision, int chunk, const issl_tuning &tn, const DevResults &out,
                   uint64_t *distinct, Workspace &ws, hipStream_t stream,
                   const Window *win = nullptr);
// The in-order form (ISSL_FORM_ORDERED): the planner, the tile kernel with
// one tile per guide and FP64 sums added hit by hit in the reference's visit
// order, and the finalizer for guides without postings.  ordered_fits: the
// index is resident and no guide can visit 2^26 - 64 postings or more.
bool ordered_fits(const issl_index *ix);
int launch_ordered(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                   const issl_tuning &tn, const DevResults &out, uint64_t *distinct,
                   Workspace &ws, hipStream_t stream);
int build_inline_layout(issl_index *ix);

// The index's scoring workspace is shared by all its calls: acquire makes a
// launch on `st` wait for the previous user when that ran on another stream;
// release records this stream as the last user (index mutex held for both).
int ws_acquire(issl_index *ix, hipStream_t st);
int ws_release(issl_index *ix, hipStream_t st);

// Parameter check and the scalars the kernels take (the checks of issl_score).
int params_to_scalars(const issl_index *ix, const issl_params *p, ScoreScalars *sc);
// One scoring launch sequence on `stream` with an explicit workspace and
// tuning; takes no lock, so callers with their own workspace and stream may
// run concurrently on one index (the index arrays are read-only).
int score_with(issl_index *ix, const uint64_t *d_guides, uint64_t n, const issl_params *p,
               const ScoreScalars &sc, const issl_tuning &tn, const DevResults &o, Workspace &ws,
               hipStream_t stream);

// Byte layout of the per-call result columns for n guides: all seven, or
// only the columns a caller asked for, packed one after another (the others
// at kNone, bound to null device pointers, so the kernels skip them).
struct OutLayout {
    static constexpr size_t kNone = ~static_cast<size_t>(0);
    size_t mit, cfd, msum, csum, occ, sites, pass, total;
    OutLayout(uint64_t n, const issl_results &want) {
        const size_t d = n * 8;
        size_t off = 0;
        auto take = [&](const void *w, size_t bytes) {
            if (!w) return kNone;
            const size_t o = off;
            off += bytes;
            return o;
        };
        mit = take(want.mit, d);
        cfd = take(want.cfd, d);
        msum = take(want.mit_sum, d);
        csum = take(want.cfd_sum, d);
        occ = take(want.occ_count, d);
        sites = take(want.site_count, d);
        pass = take(want.pass, ((n + 255) / 256) * 256);
        total = off ? off : 256;
    }
    explicit OutLayout(uint64_t n) {
        const size_t d = n * 8;
        mit = 0;
        cfd = mit + d;
        msum = cfd + d;
        csum = msum + d;
        occ = csum + d;
        sites = occ + d;
        pass = sites + d;
        total = pass + ((n + 255) / 256) * 256;
    }
    DevResults bind(char *base) const {
        auto at = [&](size_t off) { return off == kNone ? nullptr : base + off; };
        DevResults o;
        o.mit = reinterpret_cast<double *>(at(mit));
        o.cfd = reinterpret_cast<double *>(at(cfd));
        o.mit_sum = reinterpret_cast<double *>(at(msum));
        o.cfd_sum = reinterpret_cast<double *>(at(csum));
        o.occ = reinterpret_cast<uint64_t *>(at(occ));
        o.sites = reinterpret_cast<uint64_t *>(at(sites));
        o.pass = at(pass);
        return o;
    }
};
// Launch shape used before any tuning.
issl_tuning default_tuning();

// Makes `dev` current for the scope, restoring the previous device.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Index construction steps shared by the resident and streamed forms
// (device.hip): argument checks (the index's own shape first: N, slice
// count and the 20-position sequence length; then the device), slice metadata from the masks, list starts
// from the list sizes, and the device arrays (with or without the S*N
// posting arrays), plus a staged host-to-device copy and the destructor.
int check_index_args(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices, int device);
int init_meta(issl_index *ix, uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
              const uint64_t *slice_masks, uint64_t *lstart_count);
int build_lstart(issl_index *ix, const std::vector<const uint64_t *> &sizes,
                 std::vector<uint32_t> &lstart);
int alloc_index(issl_index *ix, uint64_t lstart_count, bool postings);
int upload(issl_index *ix, void *dst, const void *src, size_t bytes);
void free_index(issl_index *ix);
// issl_index_build over host arrays, or (on_device) over device arrays that
// stay the caller's.
int index_build(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices, const uint64_t *sigs,
                const uint32_t *occ, bool on_device, const uint64_t *slice_masks, int device,
                issl_index **out);

// Streamed index (stream.hip): postings stay in host memory and pass through
// a bounded set of device slice buffers.
void stream_free(issl_index *ix);
int score_streamed(issl_index *ix, const uint64_t *d_guides, uint64_t n, const issl_params *p,
                   const ScoreScalars &sc, const DevResults &o, uint64_t *distinct,
                   hipStream_t st);
// Event pair for the dominant kernel of a launch when timing is enabled
// (both null otherwise).  The pair is attached to that kernel's dispatch with
// hipExtLaunchKernelGGL, so it times the kernel itself, not the gaps around it.
void timing_events(issl_index *ix, hipEvent_t *start, hipEvent_t *stop);

// Dispatch the dominant kernel of a launch under the timing mode.  launch(e0,
// e1) issues one dispatch with those events attached (null: none).  With
// timing at N > 1 the kernel, which is idempotent, is dispatched N times back
// to back between two recorded events, so the per-dispatch average carries
// 1/N of the event overhead; results are unchanged.
template <typename F>
int timed_dispatch(issl_index *ix, hipStream_t st, F &&launch) {
    hipEvent_t e0, e1;
    timing_events(ix, &e0, &e1);
    if (ix->timing <= 1 || !e0) return launch(e0, e1);
    ISSL_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < ix->timing; ++r) {
        const int rc = launch(nullptr, nullptr);
        if (rc != ISSL_OK) return rc;
    }
    ISSL_HIP(hipEventRecord(e1, st));
    return ISSL_OK;
}

}  // namespace issl
