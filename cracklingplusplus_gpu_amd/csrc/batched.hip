// batched.hip -- BATCHED scoring form: persistent waves over (guide, chunk)
// tiles of each guide's slice-major candidate sequence.
//
// Pipeline (all on one stream, no memsets):
//   k_plan     one lane per (guide, slice): slice keys -> list ranges,
//              their prefix, the posting offsets, the tile length
//              C_g = max(chunk, ceil(P_g/64)) and the tile count K_g; the
//              guide's tiles (slot v = k*n + g) are appended to a device
//              tile queue, tiles of several-tile guides and long single
//              tiles ahead of the other single ones
//   k_tiles    persistent waves that take tiles from the queue one at a time
//              (from one of up to 256 sub-queues, so that no counter takes
//              many atomics), so waves finish together whatever a tile
//              costs (candidate-dense tiles of repeat families cost several
//              times a plain one).  For each tile it streams
//              the site words (sig | occ<<40: no signature gather), folds and
//              filters them (dist <= 4), queues the candidates with their
//              words in an LDS ring, scores them 64 at a time under the
//              slice-order first-seen rule and writes one partial per tile,
//              reduced in a fixed lane order.  The next tile's plan is loaded
//              while the current tile streams.
//   k_finalize adds the tile partials in tile order (one lane per guide when
//              no early exit is armed, else one wave per guide) and,
//              when the early exit is armed, rescans the tile in which the
//              running total first meets the exit test, hit by hit, to stop
//              exactly where the reference's loop stops
//              (lib/ISSLScoringModule.cpp:365-401)
// Results therefore depend only on the guide and the tile length, never on
// which wave ran which tile or on what else was in the batch.
//
// Limits: a launch takes at most kMaxBatch guides (larger calls run as
// several launches) and P_g < 2^31 (checked against the index's longest
// lists), so positions, tile-relative positions (< 2^26) and slots fit
// 32 bits.
//
// Precision: ISSL_PREC_FP32C forms each site's MIT/CFD product in FP32 and
// carries each lane's sums as unevaluated pairs (hi, lo) of FP32 built with
// error-free TwoSum; a lane's pair becomes an exact double, and the 64 lane
// partials, the tiles and the exit-tile rescan combine in FP64, as do the
// 10000/(100+S) transform and the threshold tests.  ISSL_PREC_FP64 uses FP64
// throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "device.hpp"

namespace issl {

// Tiles per guide at most: kMaxTiles in launches of up to kManyTilesGuides
// guides, kFewTiles in larger ones (the partials array holds n * cap
// entries).  A long guide's tiles are at least P / cap postings long.
constexpr int kMaxTiles = 256;
constexpr int kFewTiles = 64;
constexpr uint64_t kManyTilesGuides = 32768;
__host__ __device__ inline uint32_t tiles_cap(uint64_t n) {
    return n <= kManyTilesGuides ? static_cast<uint32_t>(kMaxTiles) : static_cast<uint32_t>(kFewTiles);
}
// In-order split guides have at most kSplitTiles tiles (k_ordered_terms
// holds their hit counts one per lane).
constexpr int kSplitTiles = 64;
constexpr int kTileBlock = 256;      // threads per persistent workgroup
constexpr int kTileWaves = kTileBlock / 64;

// Sums of one tile.  Both precisions combine lane partials and tiles in
// FP64; FP32C accumulates each lane's terms as compensated FP32 pairs first.
struct TilePart {
    double m, c;
    uint64_t occ;
    uint64_t sites;
};

// Guides per launch: slots n * tiles_cap(n) stay below 2^32 and the partials
// (tiles_cap(n) * n * 32 B) stay modest.
constexpr uint64_t kMaxBatch = 1ULL << 18;

// Tile queue: the planner appends to kRegions regions with a counter each,
// the tile kernel takes from up to kSubQueues sub-queues, each counter on its
// own 64 bytes (atomics on one address serialize at the memory side, ~12 ns
// apiece).
constexpr uint32_t kRegions = 32;
constexpr uint32_t kSubQueues = 256;
constexpr uint32_t kCtrStride = 16;
// The queue is kLists lists, longest tiles first: list 0 holds the tiles of
// several-tile guides and the single tiles longer than half a chunk; list j
// (1 <= j < kLists) the single tiles of (h / 2^j, h / 2^(j-1)] postings for
// h = half a chunk, the last list also the shorter ones.  Region counter
// (list j, region r) sits at tctr[(j * kRegions + r) * kCtrStride].
constexpr uint32_t kLists = 6;
constexpr uint32_t kSubQueueBase = kLists * kRegions * kCtrStride;   // taken counters
constexpr uint32_t kSubQueueCtr = kCtrStride * kSubQueues;      // taken -> finished
// In-order split: the term entries handed out (64-bit) and the split guides.
constexpr uint32_t kTermCtr = kSubQueueBase + 2 * kSubQueueCtr;
constexpr uint32_t kSplitCtr = kTermCtr + kCtrStride;
constexpr int kQueueCtrs = static_cast<int>(kSplitCtr + kCtrStride);
// Guides a region takes at most: planner block b (kPlanWaves waves of 64 / S
// guides each) appends to region b % kRegions; a region of list 0 holds
// their tiles (at most tmax per guide), of the other lists one tile per
// guide.
constexpr uint64_t kPlanWavesPerBlock = 4;
__host__ __device__ inline uint64_t region_guides(uint64_t n, int S) {
    const uint64_t G = kPlanWavesPerBlock * (64u / static_cast<uint64_t>(S > 0 ? S : 1));
    return ((n + G - 1) / G + kRegions - 1) / kRegions * G;
}

struct BatchWs {
    uint32_t *pre;       // [n][S+1] prefix of the list lengths; [S] = P_g
    uint64_t *delta;     // [n][S]   posting index = position + delta
    uint32_t *gchunk;    // [n] tile length C_g
    uint32_t *gtiles;    // [n] tile count K_g
    // [n][S] the guide's residual codes for the list of slice s (list-residual
    // layout), with s in bits 24..29; null: the signature-inline layout.
    uint32_t *gres;
    // Tile queue (null: the slot schedule).  Planner workgroup b appends the
    // slots of its guides to region b % kRegions of the list of their size
    // class (kLists), counting them in that region's counter (set back to
    // zero by k_finalize).  The queue is list 0's regions in order, then list
    // 1's, and so on: list 0 at tlist, kRegions regions of rb * tmax entries;
    // list j >= 1 after it, kRegions regions of rb entries each.  The tile kernel splits it into sub-queues, each with
    // a taken and a finished counter (kSubQueueCtr apart), which its last wave
    // sets back to zero.
    uint32_t *tlist;
    uint32_t *tctr;
    TilePart *part;      // [tmax][n] partial of tile (g, k) at slot k*n + g
    uint32_t tmax;       // tiles per guide at most in this launch: tiles_cap(n)
    // In-order split (ORDERED guides with more than the split threshold of
    // postings, while term entries last): the guide's tiles run in parallel
    // and write each hit's terms, in the guide's visit order, to entries
    // gterm[g] + k*C_g + i (i-th hit of tile k); k_ordered_terms then adds them
    // one at a time.  slist holds the split guides, counted at tctr[kSplitCtr].
    double2 *term_mc;    // [term_cap] (MIT term, CFD term) of a hit
    uint2 *term_oi;      // [term_cap] (occurrences, site id) of a hit
    uint32_t *gterm;     // [n] first entry of a split guide
    uint32_t *slist;     // [n] split guides
    uint32_t term_cap;   // entries (0: no split)
};

// Candidate ring entries: position within the tile (tiles are shorter than
// 2^26 postings for any guide with P < 2^32) and the slice index (< 64).
constexpr int kRelBits = 26;


// ------------------------------------------------------------ FP helpers

struct F2 {
    float s, c;
};

// Knuth's TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ F2 f2_add(F2 x, float t) {
    float s, e;
    two_sum(x.s, t, s, e);
    return F2{s, x.c + e};
}

__device__ __forceinline__ F2 f2_merge(F2 x, F2 y) {
    float s, e;
    two_sum(x.s, y.s, s, e);
    return F2{s, (x.c + y.c) + e};
}

__device__ __forceinline__ double f2_value(F2 x) {
    return static_cast<double>(x.s) + static_cast<double>(x.c);
}

__device__ __forceinline__ uint64_t bfold(uint64_t x) {
    return ((x >> 1) & kFoldLo) | (x & kFoldLo);
}

// Occurrence count as a scoring operand (a single 32-bit conversion).
template <typename T>
__device__ __forceinline__ T to_real(uint32_t v);
template <>
__device__ __forceinline__ float to_real<float>(uint32_t v) {
    return __uint2float_rn(v);
}
template <>
__device__ __forceinline__ double to_real<double>(uint32_t v) {
    return __uint2double_rn(v);
}

// Sentinel position for absent mismatches: its slice mask is 0 and its CFD
// factor is 1, so table reads need no predicate.
constexpr uint32_t kNoPos = 32;
constexpr uint32_t kCfdOne = 336;   // CFD table entry holding 1.0
constexpr int kCfdEntries = kCfdTable;

// Mismatch positions of a fold word (bit 2p set for each mismatching position
// p; at most four), ascending; kNoPos past the count.
__device__ __forceinline__ void mismatches(uint64_t f, uint32_t p[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[i] = f ? static_cast<uint32_t>(__builtin_ctzll(f)) >> 1 : kNoPos;
        f &= f - 1;
    }
}

// Rank of a k-subset (k = 1..4, positions ascending) in the MIT table: all
// smaller subsets first (sum of C(20, j) for 0 < j < k), then combinadic
// order (sum of C(p_i, i+1)).
__device__ __forceinline__ uint32_t mit_rank(const uint32_t p[4], int k) {
    const uint32_t c2 = p[1] * (p[1] - 1) / 2;
    const uint32_t c3 = p[2] * (p[2] - 1) * (p[2] - 2) / 6;
    const uint32_t c4 = p[3] * (p[3] - 1) * (p[3] - 2) * (p[3] - 3) / 24;
    uint32_t r = p[0];
    r += k >= 2 ? 20u + c2 : 0u;
    r += k >= 3 ? 190u + c3 : 0u;
    r += k >= 4 ? 1140u + c4 : 0u;
    return r;
}

// CFD product of a site at distance k: 1 for an exact match, else the NGG PAM
// factor times one factor per mismatch, multiplied in position order
// (lib/ISSLScoringModule.cpp:301-363).  cfd holds kCfdEntries values, the
// last one 1.0 for absent mismatches (an exact multiplication).
template <typename T>
__device__ __forceinline__ T cfd_product(uint64_t q, uint64_t site, const uint32_t p[4], int k,
                                         const T *cfd) {
    T t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t g = static_cast<uint32_t>(q >> (2 * p[i])) & 3u;
        const uint32_t o = static_cast<uint32_t>(site >> (2 * p[i])) & 3u;
        t[i] = cfd[std::min((p[i] << 4) | (g << 2) | (o ^ 3u), kCfdOne)];
    }
    const T c = cfd[320 + 10] * t[0] * t[1] * t[2] * t[3];
    return k == 0 ? T(1) : c;
}

// Slice-order first-seen rule (lib/ISSLScoringModule.cpp:252-258): a site
// found in slice sl is scored there only if no earlier slice holds it, i.e.
// every earlier slice contains one of its mismatch positions.  posmask[j] is
// the set of slices containing position j (posmask[kNoPos] = 0).
__device__ __forceinline__ bool first_seen(const uint32_t p[4], int sl, const uint64_t *posmask) {
    const uint64_t nm = posmask[p[0]] | posmask[p[1]] | posmask[p[2]] | posmask[p[3]];
    const uint64_t earlier = sl ? (~0ULL >> (64 - sl)) : 0ULL;
    return (earlier & ~nm) == 0;
}

// MIT value of a site at distance k = 1..4 from the closed form's factors,
// in its order: ((t1 * t2) * t3) * 100, t1 the product of 1 - weight over the
// mismatch positions p (kNoPos for absent ones, whose factor is 1.0), t2 and
// t3 by the count and the span p_last - p_0 (ISSLCreateIndex.cpp:235-261);
// the same bits as the MIT table entry mit_rank(p, k).
template <typename T>
__device__ __forceinline__ T mit_closed(const uint32_t p[4], int k, const double *om,
                                        const double2 *t23) {
    double t1 = om[p[0]] * om[p[1]];
    t1 = t1 * om[p[2]];
    t1 = t1 * om[p[3]];
    const uint32_t last = k >= 4 ? p[3] : k == 3 ? p[2] : k == 2 ? p[1] : p[0];
    const double2 f = t23[k * 20 + static_cast<int>(last - p[0])];
    return static_cast<T>(((t1 * f.x) * f.y) * 100.0);
}

// Fold of one site as a 20-bit mask (bit j set iff position j differs),
// from the two 32-bit halves of guide ^ site word: positions 0..15 in the
// low half, 16..19 in bits 0..7 of the high half (the bits above are the
// word's occurrence field and drop out).  Same positions as bfold.
__device__ __forceinline__ uint32_t fold20(uint32_t xl, uint32_t xh) {
    uint32_t a = ((xl >> 1) | xl) & 0x55555555u;
    a = (a | (a >> 1)) & 0x33333333u;
    a = (a | (a >> 2)) & 0x0F0F0F0Fu;
    a = (a | (a >> 4)) & 0x00FF00FFu;
    a = (a | (a >> 8)) & 0x0000FFFFu;
    uint32_t b = ((xh >> 1) | xh) & 0x55u;
    b = (b | (b >> 1)) & 0x33u;
    b = (b | (b >> 2)) & 0x0Fu;
    return a | (b << 16);
}

// Mismatch positions of a 20-bit fold, ascending; kNoPos past the count.
__device__ __forceinline__ void mismatches20(uint32_t c, uint32_t p[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[i] = c ? static_cast<uint32_t>(__builtin_ctz(c)) : kNoPos;
        c &= c - 1;
    }
}

// MIT rank terms by doubled position: rank = p0 + t[0][2 p1] + t[1][2 p2] +
// t[2][2 p3], where t[k-2][2p] = sum_{j<k} C(20, j) - 1 + C(p, k) -- the
// offsets of mit_rank -- and 0 at kNoPos, so absent mismatches add nothing.
// Built by the threads of a workgroup.
constexpr int kRankTerms = 2 * (kNoPos + 1);
__device__ __forceinline__ void build_rank_terms(uint32_t (*t)[kRankTerms], int tid, int nthreads) {
    for (int i = tid; i < 3 * kRankTerms; i += nthreads) {
        const uint32_t k = static_cast<uint32_t>(i) / kRankTerms;
        const uint32_t p = (static_cast<uint32_t>(i) % kRankTerms) / 2;
        uint32_t v = 0;
        if (p < 20) {
            if (k == 0) v = 20u + p * (p - 1) / 2;
            else if (k == 1) v = 190u + p * (p - 1) * (p - 2) / 6;
            else v = 1140u + p * (p - 1) * (p - 2) * (p - 3) / 24;
        }
        t[k][i % kRankTerms] = v;
    }
}

// posmask (kNoPos + 1 entries) for the first-seen rule, built by threads
// 0..kNoPos of a workgroup.
__device__ __forceinline__ void build_posmask(uint64_t *posmask, const uint64_t *fold, int S,
                                              int tid) {
    if (tid <= static_cast<int>(kNoPos)) {
        uint64_t m = 0;
        if (tid < static_cast<int>(kNoPos))
            for (int s = 0; s < S; ++s)
                if ((fold[s] >> (2 * tid)) & 1ULL) m |= 1ULL << s;
        posmask[tid] = m;
    }
}

__device__ __forceinline__ bool bexit(int method, double m, double c, double cut) {
    switch (method) {
    case ISSL_METHOD_MIT_AND_CFD: return m > cut && c > cut;
    case ISSL_METHOD_MIT_OR_CFD: return m > cut || c > cut;
    case ISSL_METHOD_AVG_MIT_CFD: return ((m + c) / 2.0) > cut;
    case ISSL_METHOD_MIT: return m > cut;
    default: return c > cut;
    }
}

// Value of lane (this lane ^ M): ds_swizzle's bit mode within 32-lane
// halves for M < 32 (no index arithmetic), ds_bpermute for M = 32.
template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (M < 32) {
        return static_cast<uint32_t>(
            __builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x1F | (M << 10)));
    } else {
        const int idx = static_cast<int>((__lane_id() ^ 32u) << 2);
        return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(idx, static_cast<int>(v)));
    }
}
template <int M>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v) {
    return (static_cast<uint64_t>(xor_lane<M>(static_cast<uint32_t>(v >> 32))) << 32) |
           xor_lane<M>(static_cast<uint32_t>(v));
}
template <int M>
__device__ __forceinline__ double xor_lane_f64(double v) {
    return __longlong_as_double(static_cast<long long>(
        xor_lane64<M>(static_cast<uint64_t>(__double_as_longlong(v)))));
}

// One butterfly step of the tile reduction: every lane adds its partner's
// partial (a + b and b + a agree, so all lanes hold the same sums after).
template <int M>
__device__ __forceinline__ void tile_reduce_step(double &m, double &c, uint64_t &occ) {
    m += xor_lane_f64<M>(m);
    c += xor_lane_f64<M>(c);
    occ += xor_lane64<M>(occ);
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(v), src));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(v >> 32), src));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// ------------------------------------------------------------ plan / scan

struct PlanArgs {
    const uint64_t *guides;
    uint64_t n;
    const uint32_t *lstart;
    const SliceMeta *meta;
    int S;
    uint64_t n_sites;
    uint32_t chunk;
    uint32_t first_min;           // single tiles longer than this join the first list
    int s0, s1;                   // slice window: lists outside it are empty
    const GuideState *state_in;   // guides that took the exit visit nothing
    BatchWs ws;
    // In-order split (0: off): guides with more than split_at postings get
    // tiles of split_chunk postings or longer (at most kSplitTiles) and term
    // entries, while ws.term_cap lasts.
    uint32_t split_at;
    uint32_t split_chunk;
};

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int off) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), off));
    const uint32_t hi = static_cast<uint32_t>(__shfl_up(static_cast<int>(v >> 32), off));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One lane per (guide, slice): a wave plans G = 64 / S guides, each on S
// consecutive lanes, and a workgroup kPlanWaves waves.  Every lane forms its
// slice key and requests its list's bounds (and, for the list-residual words,
// forms the guide's residual codes while they are in flight); the prefix over
// a guide's slices is a shuffle scan within its lanes.  Lists outside the
// window, or of a guide that took the exit, are loaded and then counted as
// empty.
constexpr int kPlanWaves = static_cast<int>(kPlanWavesPerBlock);
constexpr int kPlanBlock = 64 * kPlanWaves;

__host__ __device__ inline uint32_t plan_guides_per_wave(int S) {
    return 64u / static_cast<uint32_t>(S > 0 ? S : 1);
}

__global__ __launch_bounds__(kPlanBlock) void k_plan(PlanArgs a) {
    __shared__ uint32_t s_tot[kPlanWaves][kLists];
    __shared__ uint32_t s_base[kLists];
    __shared__ uint64_t s_want[kPlanWaves];
    __shared__ uint64_t s_tbase;
    const int S = a.S;
    const int SS = S > 0 ? S : 1;
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int G = static_cast<int>(plan_guides_per_wave(S));
    const int gi = lane / SS, s = lane - gi * SS;
    const uint64_t g = (blockIdx.x * static_cast<uint64_t>(kPlanWaves) + wave) * G + gi;
    const bool in = gi < G && g < a.n;
    const bool sl_in = in && s < S;
    const uint64_t q = a.guides[in ? g : 0] & kSigBits;   // n >= 1 when launched
    uint64_t fold = 0;
    uint32_t b = 0, e = 0;
    if (sl_in) {
        fold = a.meta[s].fold;
        const uint32_t lb = a.meta[s].lbase;
        const uint32_t key = slice_key(q, fold);
        b = a.lstart[lb + key];
        e = a.lstart[lb + key + 1];
    }
    // (formed while the list bounds are in flight)
    const uint32_t gr = a.ws.gres && sl_in ? residual_code_kept(q, a.meta[s].keep) | (static_cast<uint32_t>(s) << 24) : 0u;
    const bool live = in && (!a.state_in || (a.state_in[g].sites & kExited) == 0);
    const uint32_t len = (sl_in && live && s >= a.s0 && s < a.s1) ? e - b : 0u;
    // Inclusive scan of the lengths over the guide's lanes (lane - off is the
    // same guide's lane exactly when s >= off).
    uint32_t x = len;
    for (int off = 1; off < SS; off <<= 1) {
        const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), off));
        if (s >= off) x += v;
    }
    const uint32_t run = x - len;
    const uint32_t P = static_cast<uint32_t>(__shfl(static_cast<int>(x), gi * SS + SS - 1));
    if (sl_in) {
        a.ws.pre[g * (S + 1) + s] = run;
        // posting index of position p in slice s: (s - s0)*N + b + (p - pre[s])
        a.ws.delta[g * S + s] = len ? static_cast<uint64_t>(s - a.s0) * a.n_sites + b - run : 0;
        if (a.ws.gres) a.ws.gres[g * S + s] = gr;
    }
    // (P < 2^31: the launch refuses indexes whose guides could visit more,
    // so the tile arithmetic holds in 32 bits)
    const uint32_t tmax = a.ws.tmax;
    uint32_t c = (P + tmax - 1) / tmax;
    c = (c + 63u) & ~63u;
    if (c < a.chunk) c = a.chunk;
    uint32_t K = (P + c - 1) / c;
    const bool lead = in && s == 0;
    if (a.split_at) {
        // In-order split: P term entries per long guide, one 64-bit atomic per
        // workgroup; a guide whose entries would pass term_cap stays one tile.
        const uint64_t want = (lead && P > a.split_at) ? P : 0;
        uint64_t incl = want;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t v = shfl_up64(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane == 63) s_want[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0;
            for (int w = 0; w < kPlanWaves; ++w) t += s_want[w];
            s_tbase = t ? atomicAdd(reinterpret_cast<unsigned long long *>(a.ws.tctr + kTermCtr),
                                    static_cast<unsigned long long>(t))
                        : 0;
        }
        __syncthreads();
        uint64_t start = s_tbase + incl - want;
        for (int w = 0; w < wave; ++w) start += s_want[w];
        uint32_t cs = (P + kSplitTiles - 1) / kSplitTiles;
        cs = (cs + 63u) & ~63u;
        if (cs < a.split_chunk) cs = a.split_chunk;
        const uint32_t Ks = (P + cs - 1) / cs;
        if (want && Ks > 1 && start + want <= a.ws.term_cap) {
            c = cs;
            K = Ks;
            a.ws.gterm[g] = static_cast<uint32_t>(start);
            a.ws.slist[atomicAdd(a.ws.tctr + kSplitCtr, 1u)] = static_cast<uint32_t>(g);
        }
    }
    if (lead) {
        a.ws.pre[g * (S + 1) + S] = P;
        a.ws.gchunk[g] = static_cast<uint32_t>(c);
        a.ws.gtiles[g] = K;
    }
    if (a.ws.tlist) {
        // Append the workgroup's tiles to the queue: one atomic per list per
        // workgroup, on the counter of its region.  Tiles go to the list of
        // their size class, longest first (the queue's order), so that the
        // costly ones start early and the cheap ones fill the end: list 0
        // takes several-tile guides and single tiles longer than first_min
        // (half a chunk), list j the single tiles of (first_min / 2^j,
        // first_min / 2^(j-1)] postings, the last list all shorter ones.
        const bool first = K > 1 || (K == 1 && P > a.first_min);
        const bool single = lead && !first && K == 1;
        uint32_t cls = 0;
        if (!first) {
            cls = 1;
            for (uint32_t t = a.first_min >> 1; cls < kLists - 1 && P <= t; t >>= 1) ++cls;
        }
        const uint32_t ka = (lead && first) ? K : 0;
        uint32_t ia = ka;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t va = static_cast<uint32_t>(__shfl(static_cast<int>(ia), lane >= off ? lane - off : lane));
            if (lane >= off) ia += va;
        }
        // Single tiles: one per lane of the class's ballot.
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t j = 1; j < kLists; ++j) {
            const uint64_t m = __ballot(single && cls == j);
            if (single && cls == j)
                rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                 __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
            if (lane == 0) s_tot[wave][j] = static_cast<uint32_t>(__popcll(m));
        }
        if (lane == 63) s_tot[wave][0] = ia;
        __syncthreads();
        const uint32_t r = blockIdx.x % kRegions;
        if (threadIdx.x < kLists) {
            const uint32_t j = threadIdx.x;
            uint32_t t = 0;
            for (int w = 0; w < kPlanWaves; ++w) t += s_tot[w][j];
            s_base[j] = t ? atomicAdd(a.ws.tctr + (j * kRegions + r) * kCtrStride, t) : 0u;
        }
        __syncthreads();
        const uint64_t rg = region_guides(a.n, S);
        const uint32_t n32 = static_cast<uint32_t>(a.n);
        // A region holds all its blocks' tiles; the bounds only guard the
        // lists against counters that were not zero.
        {
            // list 0: each guide's tiles written by the whole wave, guide by guide
            uint32_t base = s_base[0] + ia - ka;
            for (int w = 0; w < wave; ++w) base += s_tot[w][0];
            uint32_t *const la = a.ws.tlist + r * rg * tmax;
            for (uint64_t m = __ballot(ka != 0); m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                const uint32_t kl = static_cast<uint32_t>(__shfl(static_cast<int>(ka), l));
                const uint32_t bl = static_cast<uint32_t>(__shfl(static_cast<int>(base), l));
                const uint32_t gl = static_cast<uint32_t>(__shfl(static_cast<int>(g), l));
                for (uint32_t k = static_cast<uint32_t>(lane); k < kl; k += 64)
                    if (bl + k < rg * tmax) la[bl + k] = k * n32 + gl;
            }
        }
        if (single) {
            uint32_t base = s_base[cls] + rank;
            for (int w = 0; w < wave; ++w) base += s_tot[w][cls];
            uint32_t *const lj = a.ws.tlist + kRegions * rg * tmax + ((cls - 1) * kRegions + r) * rg;
            if (base < rg) lj[base] = static_cast<uint32_t>(g);
        }
    }
}

// Final per-guide results: transform 10000/(100+S), the call and the output
// columns (lib/ISSLScoringModule.cpp:404-455).
__device__ __forceinline__ void write_guide(const ScoreScalars &sc, const DevResults &o,
                                            uint64_t g, double smit, double scfd, uint64_t occ,
                                            uint64_t sites) {
    const double vm = 10000.0 / (100.0 + smit);
    const double vc = 10000.0 / (100.0 + scfd);
    const double th = sc.threshold;
    bool reject;
    switch (sc.method) {
    case ISSL_METHOD_MIT_AND_CFD: reject = vm < th && vc < th; break;
    case ISSL_METHOD_MIT_OR_CFD: reject = vm < th || vc < th; break;
    case ISSL_METHOD_AVG_MIT_CFD: reject = (vm + vc) / 2 < th; break;
    case ISSL_METHOD_MIT: reject = vm < th; break;
    default: reject = vc < th; break;
    }
    if (o.mit) o.mit[g] = sc.calc_mit ? vm : -1.0;
    if (o.cfd) o.cfd[g] = sc.calc_cfd ? vc : -1.0;
    if (o.mit_sum) o.mit_sum[g] = smit;
    if (o.cfd_sum) o.cfd_sum[g] = scfd;
    if (o.occ) o.occ[g] = occ;
    if (o.sites) o.sites[g] = sites;
    if (o.pass) o.pass[g] = reject ? '0' : '1';
}

// ------------------------------------------------------------ tiles

struct TileArgs {
    const uint64_t *guides;
    const uint64_t *site;   // signature-inline words
    const uint32_t *res;    // list-residual words (RES kernels)
    uint32_t res_shift;     // their first occurrence bit
    const uint64_t *post;   // postings (escape occ, ids)
    const SliceMeta *meta;
    const double *mitf;     // MIT closed-form factors (kMitFactorEntries)
    const void *cfd;
    uint32_t n;             // guides; n * ws.tmax < 2^32
    int S;
    int mark_distinct;      // mark distinct sites here (no early exit armed)
    int calc_mit, calc_cfd;
    uint64_t *distinct;
    ScoreScalars sc;        // the exit test of the in-order form
    BatchWs ws;
};

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
// Candidate ring per wave: fewer than 64 pending plus kRingChk loads of 64.
constexpr int kRingChk = 3;
constexpr uint32_t kRing = 64 * (kRingChk + 1);
// Hit ring per wave: fewer than 64 pending plus one check's 64.
constexpr uint32_t kHitRing = 128;

// Postings in flight per lane: DEPTH.  QUEUE: tiles come from the planner's
// queue (else from the slot schedule).  RES: the kernel streams the
// list-residual words (4 B per posting) instead of the signature-inline words
// (8 B); candidates, hits, their order and every term are the same.
//
// In RES form a mismatch is carried as a CFD-table index 4p + x (position p,
// guide ^ site = x there), 7 bits, kAbsent7 for an absent one: s_cq[kAbsent7]
// is 1.0, and so is the MIT factor of its position (kAbsent7 / 4 = 20 lies
// past the sequence: 1 - weight is 1.0 there).  The RES kernels' factor table
// therefore holds kResCq entries per wave, the inline kernels' one entry per
// (doubled position up to kNoPos, base).
constexpr uint32_t kAbsent7 = 80;
constexpr int kResCq = 84;
static_assert(kAbsent7 / 4 == 20 && kAbsent7 < kResCq && kAbsent7 < 128, "absent code");

// Waves per SIMD the tile kernels are built for: the ORDERED list-residual
// kernels fit five (96 VGPRs, under 32 KB of LDS per workgroup with the
// table above); the ORDERED signature-inline ones are held to four by their
// LDS; the BATCHED kernels at depth 8 fit five, at other depths the compiler
// is left free.
constexpr int tile_waves_per_simd(int depth, bool queue, bool res, bool ord) {
    return ord ? (res ? 5 : 4) : (depth == 8 && queue) ? 5 : 1;
}

// Value of lane `src` (wave-uniform) of a double.
__device__ __forceinline__ double readlane_f64(double v, uint32_t src) {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b)), static_cast<int>(src)));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b >> 32)), static_cast<int>(src)));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// Value of lane `src` (wave-uniform) of a 64-bit integer, and of a float or
// double.
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t src) {
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(src)));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(src)));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ float readlane_real(float v, uint32_t src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), static_cast<int>(src)));
}
__device__ __forceinline__ double readlane_real(double v, uint32_t src) { return readlane_f64(v, src); }

// Adds p[0], p[2], p[4], ... (n terms, every other double) to `a`, one at a
// time in order, U terms read before they are added.  Lanes 0 and 1 run it
// together on the (MIT, CFD) pairs of an LDS array of double2, lane 0 from .x
// and lane 1 from .y: one FP64 addition per term and chain, where one lane
// holding both sums issues two.
template <int U>
__device__ __forceinline__ void add_chain(const double *p, uint32_t n, double &a) {
    double acc = a;
    uint32_t t = 0;
    for (; t + U <= n; t += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p[2 * (t + u)];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    for (; t < n; ++t) acc += p[2 * t];
    a = acc;
}

// add_chain that also leaves the running sum after each term in place of the
// term: p[2t] becomes a + p[0] + p[2] + ... + p[2t], added in that order.
template <int U>
__device__ __forceinline__ void add_chain_prefix(double *p, uint32_t n, double &a) {
    double acc = a;
    uint32_t t = 0;
    for (; t + U <= n; t += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p[2 * (t + u)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc += v[u];
            p[2 * (t + u)] = acc;
        }
    }
    for (; t < n; ++t) {
        acc += p[2 * t];
        p[2 * t] = acc;
    }
    a = acc;
}

// ORD: every guide is a single tile (the planner's tile length covers the
// longest guide), and each batch of hits is added to two wave-uniform FP64
// sums one hit at a time, in hit-ring order -- the guide's slice-major
// posting order -- with the early-exit test after each hit, as
// lib/ISSLScoringModule.cpp:231-401 adds and stops.  Sums, counts and the
// exit point are the reference's bit for bit; after the exit the wave stops
// streaming the guide.  (FP64 and the tile queue only.)
template <int PREC, int DEPTH, bool QUEUE, bool RES, bool ORD = false>
__global__ __launch_bounds__(kTileBlock, tile_waves_per_simd(DEPTH, QUEUE, RES, ORD)) void k_tiles(TileArgs a) {
    static_assert(!ORD || (PREC == ISSL_PREC_FP64 && QUEUE), "in-order tiles are FP64 and queued");
    using Real = typename std::conditional<PREC == ISSL_PREC_FP32C, float, double>::type;
    // Site word as streamed and ringed; the hit ring holds the word (or, in
    // RES form, the clipped occurrence count).
    using Word = typename std::conditional<RES, uint32_t, uint64_t>::type;
    __shared__ Real s_cfd[kCfdEntries];
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint64_t s_posmask[kNoPos + 1];
    // MIT closed-form factors: 1 - weight by position, (t2, t3) by count and
    // span (kMitOneMinus + 2 (20 k + span)).
    __shared__ double s_om[kMitOneMinus];
    __shared__ double2 s_t23[5 * 20];
    // Tile queue: each list's regions' exclusive prefixes, and the lists'
    // first entries (s_qbase[kLists]: the queue's length).
    __shared__ uint32_t s_qpre[QUEUE ? kLists : 1][kRegions];
    __shared__ uint32_t s_qbase[kLists + 1];
    // The current tile's CFD factors by (position, guide ^ site there).
    constexpr int CQ = RES ? kResCq : static_cast<int>(kNoPos + 1) * 4;
    __shared__ Real s_cq[kTileWaves][CQ];
    // Arrays sized by the slice count S, in dynamic LDS (tiles_lds_bytes):
    //   delta [kTileWaves][2][S]   tile plans' posting offsets, double-buffered
    //                              (the next tile's plan lands in the other half)
    //   rpm   [S][kResMaxCodes]    RES: the slices holding each residual
    //                              position (posmask by residual index), read
    //                              beside the position so stage A needs one
    //                              LDS round trip
    //   pre   [kTileWaves][2][S+1] tile plans' prefix of the list lengths
    //   gres  [kTileWaves][2][S]   RES: the guide's residual codes per slice |
    //                              slice << 24
    //   rabs  [S][kResMaxCodes]    RES: slice s's residual positions,
    //                              ascending (the code field's order; at most
    //                              kResMaxCodes, as the layout needs every
    //                              slice to have kResMinWidth or more positions)
    extern __shared__ __align__(16) char s_dyn[];
    // Candidate ring: tile-relative position | slice << kRelBits, and the site
    // word (RES: both in one 8-byte entry, written and read at once).
    __shared__ uint32_t s_rpos[kTileWaves][RES ? 1 : kRing];
    __shared__ Word s_rword[kTileWaves][RES ? 1 : kRing];
    __shared__ uint2 s_rring[kTileWaves][RES ? kRing : 1];
    // Hit ring: first-seen candidates, one 16-byte entry each: the candidate
    // ring's position entry, the packed mismatch positions and distance, and
    // the word (RES: the clipped occurrence count).
    __shared__ uint4 s_hring[kTileWaves][kHitRing];
    // ORD: a batch's hit terms (MIT, CFD) in hit order, for the in-order adds.
    __shared__ double2 s_ord[ORD ? kTileWaves : 1][ORD ? 64 : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int SD = a.S;
    uint64_t *const d_delta = reinterpret_cast<uint64_t *>(s_dyn);
    uint64_t *const d_rpm = d_delta + kTileWaves * 2 * SD;
    uint32_t *const d_pre = reinterpret_cast<uint32_t *>(d_rpm + (RES ? SD * kResMaxCodes : 0));
    uint32_t *const d_gres = d_pre + kTileWaves * 2 * (SD + 1);
    uint8_t *const d_rabs = reinterpret_cast<uint8_t *>(d_gres + (RES ? kTileWaves * 2 * SD : 0));
    // This wave's plan buffers b = 0, 1.
    auto pre_of = [&](int b) { return d_pre + (wave * 2 + b) * (SD + 1); };
    auto delta_of = [&](int b) { return d_delta + (wave * 2 + b) * SD; };
    auto gres_of = [&](int b) { return d_gres + (wave * 2 + b) * SD; };
    const Real *cfdg = static_cast<const Real *>(a.cfd);
    // Queue schedule: sub-queue sq (of nq) is served by the workgroups b with
    // b % nq == sq, sq_waves waves in all, and holds queue entries sq,
    // sq + nq, sq + 2 nq, ...  The wave of rank r among them takes entry r
    // first; every later tile comes from a ticket t (the sub-queue's taken
    // counter), entry sq_waves + t, taken when the wave has streamed its
    // current tile: the queue is ordered longest first, and a wave takes its
    // next tile only when it is nearly free, so the short tiles at the end
    // go to the waves that finish first.  The queue regions' counts are
    // requested before the tables are built, so their round trips overlap.
    const uint32_t nq = std::min<uint32_t>(kSubQueues, gridDim.x);
    const uint32_t sq = blockIdx.x % nq;
    const uint32_t sq_waves = (gridDim.x - sq + nq - 1) / nq * kTileWaves;
    const uint32_t sq_rank = (blockIdx.x / nq) * kTileWaves + static_cast<uint32_t>(wave);
    uint32_t *const sq_taken = QUEUE ? a.ws.tctr + kSubQueueBase + sq * kCtrStride : nullptr;
    uint32_t *const sq_done = QUEUE ? sq_taken + kSubQueueCtr : nullptr;
    uint32_t tk = 0;          // lane 0: the pending ticket
    uint32_t cq[kLists];      // wave 0, lane r < kRegions: region r's count in each list
#pragma unroll
    for (uint32_t j = 0; j < kLists; ++j) cq[j] = 0;
    if constexpr (QUEUE) {
        if (wave == 0 && lane < static_cast<int>(kRegions)) {
#pragma unroll
            for (uint32_t j = 0; j < kLists; ++j) cq[j] = a.ws.tctr[(j * kRegions + lane) * kCtrStride];
        }
    }
    for (int i = tid; i < kCfdEntries; i += kTileBlock) s_cfd[i] = cfdg[i];
    if (tid < a.S) s_fold[tid] = a.meta[tid].fold;
    if (tid < kMitOneMinus) s_om[tid] = a.mitf[tid];
    if (tid < 5 * 20)
        s_t23[tid] = make_double2(a.mitf[kMitOneMinus + 2 * tid], a.mitf[kMitOneMinus + 2 * tid + 1]);
    if constexpr (RES) {
        if (tid < a.S) {
            const uint64_t f = a.meta[tid].fold;
            uint32_t j = 0;
            for (uint32_t p = 0; p < static_cast<uint32_t>(kResPositions); ++p)
                if (!((f >> (2 * p)) & 1ULL) && j < static_cast<uint32_t>(kResMaxCodes))
                    d_rabs[tid * kResMaxCodes + j++] = static_cast<uint8_t>(p);
            for (; j < static_cast<uint32_t>(kResMaxCodes); ++j)
                d_rabs[tid * kResMaxCodes + j] = static_cast<uint8_t>(kNoPos);
        }
    }
    if constexpr (QUEUE) {
        // (the bounds only guard the lists against counters that were not zero)
        if (wave == 0) {
            const uint64_t rbq = region_guides(a.n, a.S);
            uint32_t base = 0;
#pragma unroll
            for (uint32_t j = 0; j < kLists; ++j) {
                const uint32_t c = static_cast<uint32_t>(std::min<uint64_t>(cq[j], j == 0 ? rbq * a.ws.tmax : rbq));
                uint32_t x = c;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(x), lane >= off ? lane - off : lane));
                    if (lane >= off) x += v;
                }
                if (lane < static_cast<int>(kRegions)) s_qpre[j][lane] = x - c;
                if (lane == 0) s_qbase[j] = base;
                base += static_cast<uint32_t>(__shfl(static_cast<int>(x), 63));
            }
            if (lane == 0) s_qbase[kLists] = base;
        }
    }
    __syncthreads();
    build_posmask(s_posmask, s_fold, a.S, tid);
    __syncthreads();
    if constexpr (RES) {
        for (int i = tid; i < a.S * kResMaxCodes; i += kTileBlock) {
            const uint32_t p = d_rabs[i];
            d_rpm[i] = p < kNoPos ? s_posmask[p] : 0ULL;
        }
        __syncthreads();
    }
    // CFD factors of absent mismatches (positions 20..kNoPos): 1.0 in every
    // tile (the tiles rewrite positions 0..19 only).
    for (int i = 20 * 4 + lane; i < CQ; i += 64) s_cq[wave][i] = Real(1);
    // RES: the code field's fold bits (bit 2j per residual position j) and the
    // occurrence escape.
    const uint32_t res_fold = RES ? (0x55555555u & ((1u << a.res_shift) - 1u)) : 0u;
    const uint32_t res_esc = RES ? (0xFFFFFFFFu >> a.res_shift) : 0u;
    const Real cfd_pam = s_cfd[320 + 10];   // NGG
    const int S = a.S;
    const uint32_t n = a.n;
    const uint32_t V = n * a.ws.tmax;
    const uint32_t W = gridDim.x * kTileWaves;
    const uint32_t w = blockIdx.x * kTileWaves + wave;

    // Queue schedule: a wave asks for its next tile when it has streamed
    // the current one, so that the waves finish together whatever each tile
    // costs; the queue holds the tiles longest first (kLists size classes).
    // The request (ticket, then the queue entry, then the plan) overlaps the
    // tile's last candidates and its reduction.  Slot schedule (guide streams
    // wrote the plans): the slots of this wave are w + j*W; 64 of them are
    // checked per load and the real ones (k < K_g) are taken in increasing
    // order.
    const uint64_t rb = QUEUE ? region_guides(n, a.S) : 0;
    // Queue entry c: its list is the last one starting at or before c, its
    // region the last of that list starting at or before c.  Where its slot
    // is stored, or null past the queue's end.
    auto entry_at = [&](uint64_t c) -> const uint32_t * {
        if constexpr (QUEUE) {
            if (c >= s_qbase[kLists]) return nullptr;
            uint32_t j = 0;
            while (j + 1 < kLists && s_qbase[j + 1] <= c) ++j;
            const uint32_t cj = static_cast<uint32_t>(c) - s_qbase[j];
            const uint32_t pv = lane < static_cast<int>(kRegions) ? s_qpre[j][lane] : 0xFFFFFFFFu;
            const uint64_t m = __ballot(pv <= cj);
            const int r = 63 - __builtin_clzll(m);
            const uint32_t off = cj - static_cast<uint32_t>(__shfl(static_cast<int>(pv), r));
            if (j == 0) return a.ws.tlist + r * rb * a.ws.tmax + off;
            return a.ws.tlist + kRegions * rb * a.ws.tmax + ((j - 1) * kRegions + r) * rb + off;
        }
        return nullptr;
    };
    // A wave leaving the kernel (with no ticket pending): the last one of its
    // sub-queue's waves sets the sub-queue's counters back to zero for the
    // next dispatch.
    auto leave = [&]() {
        if constexpr (QUEUE) {
            if (lane == 0 && atomicAdd(sq_done, 1u) == sq_waves - 1) {
                atomicExch(sq_taken, 0u);
                atomicExch(sq_done, 0u);
            }
        }
    };
    uint32_t j0 = 0, vl = 0;
    uint64_t smask = 0;
    // Slot schedule: the next slot of this wave.
    auto next_slot = [&]() -> uint32_t {
        while (smask == 0) {
            if (static_cast<uint64_t>(j0) * W + w >= V) return kNoSlot;
            const uint64_t vv = static_cast<uint64_t>(j0 + lane) * W + w;
            bool real = false;
            if (vv < V) {
                vl = static_cast<uint32_t>(vv);
                real = vl / n < a.ws.gtiles[vl % n];
            }
            smask = __ballot(real);
            j0 += 64;
        }
        const int l = __builtin_ctzll(smask);
        smask &= smask - 1;
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vl), l));
    };

    // A tile's plan: guide, tile index, length, and per slice (lane) the
    // prefix and the posting offset.
    struct Plan {
        uint64_t q;
        uint32_t g, k, P, C, K;
        uint32_t pre;
        uint64_t dl;
        uint32_t gr;   // RES: the guide's residual codes for slice `lane` | lane << 24
        uint32_t tb;   // ORD, split guide: the tile's first term entry
    };
    auto plan_load = [&](uint32_t v, Plan &m) {
        m.g = v % n;
        m.k = v / n;
        const size_t row = static_cast<size_t>(m.g) * (S + 1);
        m.q = a.guides[m.g] & kSigBits;
        m.P = a.ws.pre[row + S];
        m.C = a.ws.gchunk[m.g];
        m.K = a.ws.gtiles[m.g];
        m.pre = 0;
        m.dl = 0;
        m.gr = 0;
        m.tb = 0;
        // (gterm is written for split guides only, and read for K > 1 only)
        if (ORD && a.ws.term_cap) m.tb = a.ws.gterm[m.g] + m.k * m.C;
        if (lane < S) {
            m.pre = a.ws.pre[row + lane];
            m.dl = a.ws.delta[static_cast<size_t>(m.g) * S + lane];
            if constexpr (RES) m.gr = a.ws.gres[static_cast<size_t>(m.g) * S + lane];
        }
    };
    auto plan_commit = [&](const Plan &m, int b) {
        if (lane < S) {
            pre_of(b)[lane] = m.pre;
            delta_of(b)[lane] = m.dl;
            if constexpr (RES) gres_of(b)[lane] = m.gr;
        }
        if (lane == 0) pre_of(b)[S] = m.P;  // S may be 64: no lane S
        __builtin_amdgcn_wave_barrier();
    };

    // The next tile's request: stage 0, the ticket tk is in flight; 1, the
    // queue entry tv is; 2, the next slot vn is known and its plan mn is in
    // flight (or vn is kNoSlot, or no request was made).
    Plan mc, mn;
    uint32_t vc = kNoSlot, vn = kNoSlot, tv = kNoSlot;
    int stage = 2;
    auto request_step = [&]() {
        if constexpr (QUEUE) {
            if (stage == 0) {
                const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(tk)));
                const uint32_t *const e = entry_at(sq + static_cast<uint64_t>(sq_waves + t) * nq);
                tv = e ? *e : kNoSlot;
                stage = 1;
            } else if (stage == 1) {
                // (the entry is the same in every lane: keep it scalar)
                const uint32_t v = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(tv)));
                vn = v < V ? v : kNoSlot;   // a slot outside the slot space ends the wave
                if (vn != kNoSlot) plan_load(vn, mn);
                stage = 2;
            }
        }
    };
    if constexpr (QUEUE) {
        // The first tile is queue entry sq_rank.
        const uint32_t *const e = entry_at(sq + static_cast<uint64_t>(sq_rank) * nq);
        const uint32_t v = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(e ? *e : kNoSlot)));
        vc = v < V ? v : kNoSlot;
    } else {
        vc = next_slot();
    }
    if (vc == kNoSlot) {  // uniform per wave; no workgroup barrier follows
        leave();
        return;
    }
    plan_load(vc, mc);
    plan_commit(mc, 0);
    if constexpr (!QUEUE) {
        vn = next_slot();
        if (vn != kNoSlot) plan_load(vn, mn);
    }
    int buf = 0;

    for (;;) {
        const uint64_t q = mc.q;
        const uint32_t a0 = mc.k * mc.C;
        const uint32_t a1 = std::min(mc.P, a0 + mc.C);
        // Entry 4p + x: the factor of a mismatch at position p where guide ^
        // site = x (lib/ISSLScoringModule.cpp:308-361), for p < 20 (the
        // entries of positions 20..kNoPos hold 1.0 from the kernel's start).
        for (int i = lane; i < 20 * 4; i += 64) {
            const uint32_t pp = static_cast<uint32_t>(i) >> 2, x = static_cast<uint32_t>(i) & 3u;
            const uint32_t g = static_cast<uint32_t>(q >> (2 * pp)) & 3u;
            s_cq[wave][i] = s_cfd[(pp << 4) | (g << 2) | ((g ^ x) ^ 3u)];
        }
        __builtin_amdgcn_wave_barrier();

        Real am = 0, ac = 0;
        F2 fm{0.f, 0.f}, fc{0.f, 0.f};
        uint64_t occ_sum = 0;
        uint32_t sites = 0;   // hits of this tile (wave-uniform; < 2^26)
        // ORD: the guide's running sums (wave-uniform) and whether the exit
        // has been taken (then nothing more of the guide is scored).  A tile
        // of a split guide (K > 1) writes its hits' terms instead, and its
        // hit count goes to the partial.
        double run_m = 0.0, run_c = 0.0;
        bool stop = false;
        const bool tmode = ORD && mc.K > 1;
        uint32_t hcnt = 0;   // tmode: hits of the tile so far

        // Candidates (dist <= 4) wait in the ring in position order.  They are
        // checked 64 at a time against the first-seen rule, one per lane; the
        // sites that count are compacted into the hit ring with their
        // mismatch positions, and hits are scored 64 at a time.  Both rare
        // paths thus run with every lane busy.
        uint32_t head = 0, tail = 0;    // candidate ring cursors (wave-uniform)
        uint32_t hhead = 0, htail = 0;  // hit ring cursors
        auto score_hits = [&](uint32_t cnt) {
            __builtin_amdgcn_wave_barrier();
            if constexpr (ORD) {
                if (stop) {   // past the exit: nothing more counts
                    hhead += cnt;
                    return;
                }
            } else {
                sites += cnt;
            }
            uint32_t h_occ = 0, h_id = 0;
            Real h_tm = 0, h_tc = 0;
            if (static_cast<uint32_t>(lane) < cnt) {
                const uint32_t e = (hhead + lane) & (kHitRing - 1);
                const uint4 he = s_hring[wave][e];
                const uint32_t ent = he.x;
                const uint32_t pk = he.y;
                Word wv;
                if constexpr (RES) wv = he.z;
                else wv = (static_cast<uint64_t>(he.w) << 32) | he.z;
                // 2p (or, RES, 4p + x) of each mismatch
                const uint32_t m2[4] = {pk & 127u, (pk >> 7) & 127u, (pk >> 14) & 127u,
                                        (pk >> 21) & 127u};
                const int dist = static_cast<int>(pk >> 28);
                uint32_t occ;
                bool esc;
                if constexpr (RES) {
                    occ = wv;
                    esc = occ == res_esc;
                } else {
                    occ = static_cast<uint32_t>(wv >> 40);
                    esc = occ == kOccEscape;
                }
                if (esc || a.mark_distinct) {
                    const int sl = static_cast<int>(ent >> kRelBits);
                    const uint64_t addr =
                        a0 + (ent & ((1u << kRelBits) - 1)) + delta_of(buf)[sl];
                    const uint64_t pe = a.post[addr];
                    occ = static_cast<uint32_t>(pe >> 32);
                    h_id = static_cast<uint32_t>(pe);
                    if (!ORD && a.mark_distinct)
                        atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (h_id >> 6),
                                 1ULL << (h_id & 63));
                }
                if constexpr (!ORD) occ_sum += occ;
                h_occ = occ;
                // Keep occ a 32-bit value: otherwise the compiler widens it
                // back to the 64-bit word it was cut from and converts that.
                asm("" : "+v"(occ));
                const Real ro = to_real<Real>(occ);
                Real tm = 0, tc = 0, mv = 0;
                if (a.calc_mit) {
                    // The MIT value from the closed form's factors, in its
                    // order: ((t1 * t2) * t3) * 100 with t1 the product of
                    // 1 - weight over the mismatch positions (absent ones,
                    // position 31 or 32, give 1.0), t2 and t3 by the count
                    // and the span of the positions
                    // (ISSLCreateIndex.cpp:235-261); the same bits as the
                    // table the reference loads.  (An exact match reads the
                    // k = 0 entry, which is not used.)
                    constexpr int SH = RES ? 2 : 1;
                    const uint32_t p0 = m2[0] >> SH, p1 = m2[1] >> SH, p2 = m2[2] >> SH, p3 = m2[3] >> SH;
                    double t1 = s_om[p0] * s_om[p1];
                    t1 = t1 * s_om[p2];
                    t1 = t1 * s_om[p3];
                    const uint32_t last = dist >= 4 ? p3 : dist == 3 ? p2 : dist == 2 ? p1 : p0;
                    const double2 f = s_t23[dist * 20 + static_cast<int>(last - p0)];
                    mv = static_cast<Real>(((t1 * f.x) * f.y) * 100.0);
                }
                if (a.calc_cfd) {
                    // PAM factor, then one factor per mismatch in position
                    // order (1 for absent ones): the order of cfd_product.
                    Real c = cfd_pam;
                    if constexpr (RES) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) c = c * s_cq[wave][m2[i]];
                    } else {
                        const uint64_t xw = q ^ (wv & kSigBits);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const uint32_t xv = static_cast<uint32_t>(xw >> (m2[i] & 63u)) & 3u;
                            c = c * s_cq[wave][m2[i] * 2 + xv];
                        }
                    }
                    tc = (dist == 0 ? Real(1) : c) * ro;
                }
                if (a.calc_mit && dist > 0) tm = mv * ro;
                if constexpr (ORD) {
                    h_tm = tm;
                    h_tc = tc;
                } else if constexpr (PREC == ISSL_PREC_FP32C) {
                    fm = f2_add(fm, tm);
                    fc = f2_add(fc, tc);
                } else {
                    am += tm;
                    ac += tc;
                }
            }
            if (ORD && tmode) {
                // Split guide: the terms in hit order, for k_ordered_terms.
                if (static_cast<uint32_t>(lane) < cnt) {
                    const uint32_t i = mc.tb + hcnt + static_cast<uint32_t>(lane);
                    a.ws.term_mc[i] = make_double2(static_cast<double>(h_tm), static_cast<double>(h_tc));
                    a.ws.term_oi[i] = make_uint2(h_occ, h_id);
                    occ_sum += h_occ;
                }
                sites += cnt;
                hcnt += cnt;
            } else if constexpr (ORD) {
                // The hits in order, one at a time: hit i's terms sit at
                // s_ord[i], and lane 0 adds them to the running sums (the
                // reads run ahead of the dependent adds); the sums and the
                // exit point are then broadcast to the wave.
                if (static_cast<uint32_t>(lane) < cnt)
                    s_ord[wave][lane] = make_double2(static_cast<double>(h_tm), static_cast<double>(h_tc));
                __builtin_amdgcn_wave_barrier();
                uint32_t acc = cnt, ex = 0;
                if (a.sc.use_cut) {
                    // Lanes 0 and 1 leave the running MIT and CFD sums after
                    // each hit in place of its terms; lane i then tests the
                    // exit on the sums after hit i, and the first hit that
                    // passes is where the one-at-a-time loop stops, with
                    // those sums.
                    double *const pd = reinterpret_cast<double *>(s_ord[wave]);
                    double ch = lane == 1 ? run_c : run_m;
                    if (lane < 2) add_chain_prefix<8>(pd + lane, cnt, ch);
                    __builtin_amdgcn_wave_barrier();
                    bool passes = false;
                    if (static_cast<uint32_t>(lane) < cnt)
                        passes = bexit(a.sc.method, pd[2 * lane], pd[2 * lane + 1], a.sc.cut);
                    const uint64_t bx = __ballot(passes);
                    if (bx) {
                        const uint32_t f = static_cast<uint32_t>(__builtin_ctzll(bx));
                        acc = f + 1;
                        ex = 1;
                        run_m = pd[2 * f];
                        run_c = pd[2 * f + 1];
                    } else {
                        run_m = readlane_f64(ch, 0);
                        run_c = readlane_f64(ch, 1);
                    }
                } else {
                    // No exit: lane 0 adds the MIT terms and lane 1 the CFD
                    // terms, one addition per hit each (add_chain).
                    double ch = lane == 1 ? run_c : run_m;
                    if (lane < 2) add_chain<8>(reinterpret_cast<const double *>(s_ord[wave]) + lane, cnt, ch);
                    __builtin_amdgcn_wave_barrier();
                    run_m = readlane_f64(ch, 0);
                    run_c = readlane_f64(ch, 1);
                }
                acc = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(acc)));
                stop = __builtin_amdgcn_readfirstlane(static_cast<int>(ex)) != 0;
                // Counts and distinct sites of the hits up to the exit.
                sites += acc;
                if (static_cast<uint32_t>(lane) < acc) {
                    occ_sum += h_occ;
                    if (a.mark_distinct)
                        atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (h_id >> 6),
                                 1ULL << (h_id & 63));
                }
            }
            hhead += cnt;
            __builtin_amdgcn_wave_barrier();
        };
        auto drain = [&](uint32_t cnt) {
            __builtin_amdgcn_wave_barrier();
            if constexpr (ORD) {
                if (stop) {   // past the exit: the candidates are not scored
                    head += cnt;
                    return;
                }
            }
            bool hit = false;
            uint32_t ent = 0, pk = 0;
            Word wv = 0;
            if (static_cast<uint32_t>(lane) < cnt) {
                const uint32_t e = (head + lane) & (kRing - 1);
                if constexpr (RES) {
                    const uint2 r = s_rring[wave][e];
                    ent = r.x;
                    wv = r.y;
                } else {
                    ent = s_rpos[wave][e];
                    wv = s_rword[wave][e];
                }
                const int sl = static_cast<int>(ent >> kRelBits);
                if constexpr (RES) {
                    // Residual fold: bit 2j for each differing residual
                    // position j; positions ascend with j.  The ring holds
                    // x = word ^ (guide codes | slice << 24).
                    // The four lowest set bits are found first and both
                    // tables read for all four at once (an absent one reads
                    // entry kResMaxCodes - 1, via the odd bit 2 kResMaxCodes - 1,
                    // which no fold sets), so the lookups share one LDS
                    // round trip.
                    const uint32_t x = wv;
                    uint32_t fi[4];
                    fi[0] = ((x >> 1) | x) & res_fold;
#pragma unroll
                    for (int i = 1; i < 4; ++i) fi[i] = fi[i - 1] & (fi[i - 1] - 1u);
                    const uint32_t dist = static_cast<uint32_t>(__popc(fi[0]));
                    uint32_t b2[4], p[4];
                    uint64_t pm[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        b2[i] = static_cast<uint32_t>(__builtin_ctz(fi[i] | (1u << (2 * kResMaxCodes - 1))));
                    const uint8_t *const rab = d_rabs + sl * kResMaxCodes;
                    const uint64_t *const rpm = d_rpm + sl * kResMaxCodes;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        p[i] = rab[b2[i] >> 1];
                        pm[i] = rpm[b2[i] >> 1];
                    }
                    uint32_t ix[4];
                    uint64_t nm = 0;   // slices holding a mismatch position
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        ix[i] = fi[i] ? ((p[i] << 2) | ((x >> b2[i]) & 3u)) : kAbsent7;
                        nm |= fi[i] ? pm[i] : 0ULL;
                    }
                    // the first-seen rule (first_seen) on those slices
                    const uint64_t earlier = sl ? (~0ULL >> (64 - sl)) : 0ULL;
                    hit = (earlier & ~nm) == 0;
                    pk = ix[0] | (ix[1] << 7) | (ix[2] << 14) | (ix[3] << 21) | (dist << 28);
                    // the hit ring keeps the occurrence field
                    wv = (x ^ (static_cast<uint32_t>(sl) << 24)) >> a.res_shift;
                } else {
                    const uint32_t f = fold20(static_cast<uint32_t>(q ^ wv),
                                              static_cast<uint32_t>((q ^ wv) >> 32));
                    const uint32_t dist = static_cast<uint32_t>(__popc(f));
                    uint32_t mp[4];
                    mismatches20(f, mp);
                    hit = first_seen(mp, sl, s_posmask);
                    // doubled positions (the shift that picks a base), 7 bits each
                    pk = (mp[0] << 1) | (mp[1] << 8) | (mp[2] << 15) | (mp[3] << 22) | (dist << 28);
                }
            }
            head += cnt;
            const uint64_t b = __ballot(hit);
            if (b) {
                if (hit) {
                    const uint32_t e = __builtin_amdgcn_mbcnt_hi(
                        static_cast<uint32_t>(b >> 32),
                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), htail)) & (kHitRing - 1);
                    if constexpr (RES) s_hring[wave][e] = make_uint4(ent, pk, wv, 0u);
                    else
                        s_hring[wave][e] = make_uint4(ent, pk, static_cast<uint32_t>(wv),
                                                      static_cast<uint32_t>(wv >> 32));
                }
                htail += static_cast<uint32_t>(__popcll(b));
            }
            __builtin_amdgcn_wave_barrier();
            if (htail - hhead >= 64) score_hits(64);
        };

        // Per-lane slice cursor: the next slice boundary and this slice's
        // posting offset stay in registers; LDS is read only at boundaries.
        int cur = 0;
        uint32_t nb;
        uint64_t dl;
        const uint32_t *const t_pre = pre_of(buf);
        const uint64_t *const t_delta = delta_of(buf);
        const uint32_t *const t_gres = gres_of(buf);
        {
            const uint32_t p0 = a0 + lane;
            while (cur < S - 1 && p0 >= t_pre[cur + 1]) ++cur;
            nb = t_pre[cur + 1];
            dl = t_delta[cur];
        }
        // The filter works on 32-bit halves: positions 0..15 in the low word,
        // 16..19 in bits 0..7 of the high word; the mask drops the occurrence
        // bits above them and the bit carried across the halves by the shift.
        const uint32_t ql = static_cast<uint32_t>(q), qh = static_cast<uint32_t>(q >> 32);
        const uint64_t *site_p = a.site + dl;   // this lane's current slice, by position
        // RES: the same for the residual words, and the guide's residual codes
        // for the current slice | slice << 24.
        const uint32_t *res_p = RES ? a.res + dl : nullptr;
        uint32_t gsl = RES ? t_gres[cur] : 0u;
        // One round: DEPTH loads of 64 consecutive postings, then the filter.
        // FULL rounds lie inside the tile; in the last round lanes past the
        // tile read its last posting again (every lane loads without a
        // branch) and fail the filter.
        auto round = [&](uint32_t base, auto full_c) {
            constexpr bool FULL = decltype(full_c)::value;
            Word wv[DEPTH];
            uint32_t sl[DEPTH];   // slice (RES: gsl)
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const uint32_t pos = base + 64 * d + lane;
                const uint32_t pc = FULL ? pos : std::min(pos, a1 - 1);
                if (pc >= nb) {
                    do {
                        ++cur;
                        nb = t_pre[cur + 1];
                    } while (pc >= nb);
                    if constexpr (RES) {
                        res_p = a.res + t_delta[cur];
                        gsl = t_gres[cur];
                    } else {
                        site_p = a.site + t_delta[cur];
                    }
                }
                if constexpr (RES) {
                    sl[d] = gsl;
                    wv[d] = res_p[pc];
                } else {
                    sl[d] = static_cast<uint32_t>(cur);
                    wv[d] = site_p[pc];
                }
            }
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                bool cand;
                if constexpr (RES) {
                    // 8 added to the count of a lane past the tile: no candidate.
                    // The ring keeps word ^ (guide codes | slice << 24).
                    const uint32_t kill = (FULL || base + 64 * d + lane < a1) ? 0u : 8u;
                    wv[d] ^= sl[d];
                    const uint32_t f = ((wv[d] >> 1) | wv[d]) & res_fold;
                    cand = __popc(f) + kill <= 4;
                } else {
                    // All ones in the low half: 16 differing positions, no candidate.
                    const uint32_t kill = (FULL || base + 64 * d + lane < a1) ? 0u : ~0u;
                    const uint32_t xl = (ql ^ static_cast<uint32_t>(wv[d])) | kill;
                    const uint32_t xh = qh ^ static_cast<uint32_t>(wv[d] >> 32);
                    const uint32_t fl = ((xl >> 1) | xl) & 0x55555555u;
                    const uint32_t fh = ((xh >> 1) | xh) & 0x55u;
                    cand = __popc(fl) + __popc(fh) <= 4;
                }
                const uint64_t b = __ballot(cand);
                if (b) {
                    if (cand) {
                        const uint32_t e = __builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(b >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), tail)) & (kRing - 1);
                        const uint32_t slice = RES ? (sl[d] >> 24) : sl[d];
                        const uint32_t rp = (base - a0 + 64 * d + lane) | (slice << kRelBits);
                        if constexpr (RES) {
                            s_rring[wave][e] = make_uint2(rp, wv[d]);
                        } else {
                            s_rpos[wave][e] = rp;
                            s_rword[wave][e] = wv[d];
                        }
                    }
                    tail += static_cast<uint32_t>(__popcll(b));
                }
                if ((d + 1) % kRingChk == 0 || d == DEPTH - 1)
                    while (tail - head >= 64) drain(64);
            }
        };
        uint32_t base = a0;
        for (; base + 64 * DEPTH <= a1; base += 64 * DEPTH) {
            if (ORD && stop) break;
            round(base, std::true_type{});
        }
        if (base < a1 && !(ORD && stop)) round(base, std::false_type{});
        if constexpr (QUEUE) {
            // the ticket for the next tile
            if (lane == 0) tk = atomicAdd(sq_taken, 1u);
            stage = 0;
        }
        while (tail != head) drain(std::min<uint32_t>(64, tail - head));
        if (htail != hhead) score_hits(htail - hhead);
        request_step();   // the ticket's answer, and the queue entry's request
        // Lane partials as doubles (a compensated pair converts exactly),
        // combined in a fixed butterfly: every lane ends with the same sums.
        double dm, dc;
        if constexpr (ORD) {
            dm = 0.0;   // the sums are run_m, run_c; the butterfly reduces the counts
            dc = 0.0;
        } else if constexpr (PREC == ISSL_PREC_FP32C) {
            dm = static_cast<double>(fm.s) + static_cast<double>(fm.c);
            dc = static_cast<double>(fc.s) + static_cast<double>(fc.c);
        } else {
            dm = am;
            dc = ac;
        }
        tile_reduce_step<32>(dm, dc, occ_sum);
        tile_reduce_step<16>(dm, dc, occ_sum);
        tile_reduce_step<8>(dm, dc, occ_sum);
        tile_reduce_step<4>(dm, dc, occ_sum);
        tile_reduce_step<2>(dm, dc, occ_sum);
        tile_reduce_step<1>(dm, dc, occ_sum);
        if constexpr (ORD) {
            dm = run_m;
            dc = run_c;
        }
        request_step();   // the entry, and the plan's request
        if (lane == 0) {
            TilePart tp;
            tp.m = dm;
            tp.c = dc;
            tp.occ = occ_sum;
            tp.sites = sites;
            a.ws.part[vc] = tp;
        }
        if (vn == kNoSlot) {
            leave();
            break;
        }
        buf ^= 1;
        plan_commit(mn, buf);
        vc = vn;
        mc = mn;
        if constexpr (QUEUE) {
            vn = kNoSlot;
        } else {
            vn = next_slot();
            if (vn != kNoSlot) plan_load(vn, mn);
        }
    }
}

// ------------------------------------------------------------ finalize

struct FinArgs {
    const uint64_t *guides;
    const uint64_t *site;
    const uint64_t *post;
    const SliceMeta *meta;
    const void *mit;
    const void *cfd;
    const double *mitf;           // MIT closed-form factors (kMitFactorEntries)
    int S;
    uint64_t n;
    ScoreScalars sc;
    DevResults out;
    uint64_t *distinct;           // mark sites up to the exit (early exit armed)
    int skip_untiled;             // guides without tiles were finished by k_stream
    int skip_multi;               // guides with several tiles are split guides of the
                                  // in-order form, finished by k_ordered_terms
    const GuideState *state_in;   // carried state (null: start from zero)
    GuideState *state_out;        // null: write the results
    BatchWs ws;
};

// The tile queue's counters back to zero for the next launch sequence (the
// tile kernel has finished with them).
__device__ __forceinline__ void reset_queue(const BatchWs &ws) {
    if (ws.tctr && blockIdx.x == 0 && blockIdx.y == 0)
        for (uint32_t i = threadIdx.x; i < kLists * kRegions; i += blockDim.x) ws.tctr[i * kCtrStride] = 0;
    if (ws.tctr && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ws.tctr[kTermCtr] = 0;
        ws.tctr[kTermCtr + 1] = 0;
        ws.tctr[kSplitCtr] = 0;
    }
}

__global__ void k_zero_queue(uint32_t *tctr) {
    for (int i = static_cast<int>(threadIdx.x); i < kQueueCtrs; i += static_cast<int>(blockDim.x))
        tctr[i] = 0;
}

// Rows of 64 postings whose site words the exit rescan loads together.
constexpr int kRescanRows = 8;

template <int PREC>
__global__ __launch_bounds__(256) void k_finalize(FinArgs a) {
    using Real = typename std::conditional<PREC == ISSL_PREC_FP32C, float, double>::type;
    constexpr int W = 256 / 64;
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint64_t s_posmask[kNoPos + 1];
    __shared__ uint64_t s_pre[W][kMaxSlices + 1];
    __shared__ uint64_t s_delta[W][kMaxSlices];
    // The rescan's term tables: CFD factors, and the MIT closed form's
    // factors (1 - weight by position, kNoPos giving 1.0; then the count and
    // span factors as pairs).
    __shared__ Real s_cfd[kCfdEntries];
    __shared__ double s_om[kMitOneMinus];
    __shared__ double2 s_t23[5 * 20];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    reset_queue(a.ws);
    if (tid < a.S) s_fold[tid] = a.meta[tid].fold;
    for (int i = tid; i < kCfdEntries; i += 256) s_cfd[i] = static_cast<const Real *>(a.cfd)[i];
    if (tid < kMitOneMinus) s_om[tid] = a.mitf[tid];
    if (tid < 5 * 20)
        s_t23[tid] = make_double2(a.mitf[kMitOneMinus + 2 * tid], a.mitf[kMitOneMinus + 2 * tid + 1]);
    __syncthreads();
    build_posmask(s_posmask, s_fold, a.S, tid);
    __syncthreads();
    const uint64_t g = blockIdx.x * static_cast<uint64_t>(W) + wave;
    if (g >= a.n) return;
    const int S = a.S;
    const uint32_t K = a.ws.gtiles[g];
    if (K == 0 && a.skip_untiled) return;
    // Lane l holds the partial of tile k0 + l, 64 tiles at a time.
    TilePart mine{};
    uint32_t k0 = 0;
    if (lane < static_cast<int>(K)) mine = a.ws.part[static_cast<uint64_t>(lane) * a.n + g];

    // Sums so far (FP64): zero, or the state carried from earlier slice
    // windows.
    double dm = 0, dc = 0;
    uint64_t occ_sum = 0, sites = 0;
    bool exited = false;
    if (a.state_in) {
        const GuideState st = a.state_in[g];
        dm = st.m;
        dc = st.c;
        occ_sum = st.occ;
        sites = st.sites & ~kExited;
        exited = (st.sites & kExited) != 0;
    }

    const uint64_t q = a.guides[g] & kSigBits;
    const uint64_t C = a.ws.gchunk[g];
    bool lists = false;
    auto load_lists = [&]() {  // the guide's list prefix and posting offsets
        if (lists) return;
        if (lane < S) s_pre[wave][lane] = a.ws.pre[g * (S + 1) + lane];
        if (lane == 0) s_pre[wave][S] = a.ws.pre[g * (S + 1) + S];  // S may be 64
        if (lane < S) s_delta[wave][lane] = a.ws.delta[g * S + lane];
        __builtin_amdgcn_wave_barrier();
        lists = true;
    };

    int exit_tile = -1;
    for (uint32_t k = 0; k < K; ++k) {
        if (k - k0 == 64) {
            k0 = k;
            mine = TilePart{};
            if (k0 + lane < K) mine = a.ws.part[static_cast<uint64_t>(k0 + lane) * a.n + g];
        }
        const uint32_t kl = k - k0;   // wave-uniform
        const uint64_t o = readlane_u64(mine.occ, kl);
        const uint64_t si = readlane_u64(mine.sites, kl);
        const double qm = dm, qc = dc;
        dm += readlane_f64(mine.m, kl);
        dc += readlane_f64(mine.c, kl);
        if (!(a.sc.use_cut && bexit(a.sc.method, dm, dc, a.sc.cut))) {
            occ_sum += o;
            sites += si;
            continue;
        }
        // The total meets the exit test within tile k: restore the sums before
        // it and walk its postings in order, 64 at a time.  A row whose hits'
        // sum (a butterfly over the wave) leaves the test unmet is taken
        // whole; in the row where the test would be met the hits are added
        // one at a time, in order, to stop where the reference's loop stops
        // (lib/ISSLScoringModule.cpp:365-401).  The site words of
        // kRescanRows rows are loaded together.
        dm = qm;
        dc = qc;
        load_lists();
        const uint64_t P = s_pre[wave][S];
        const uint64_t a0 = static_cast<uint64_t>(k) * C;
        const uint64_t a1 = std::min<uint64_t>(P, a0 + C);
        int cur = 0;
        bool done = false;
        for (uint64_t r0 = a0; r0 < a1 && !done; r0 += 64 * kRescanRows) {
            uint64_t w[kRescanRows];
            int cs[kRescanRows];
#pragma unroll
            for (int r = 0; r < kRescanRows; ++r) {
                const uint64_t pos = r0 + 64 * r + lane;
                w[r] = 0;
                cs[r] = -1;
                if (pos < a1) {
                    while (pos >= s_pre[wave][cur + 1]) ++cur;
                    cs[r] = cur;
                    w[r] = a.site[pos + s_delta[wave][cur]];
                }
            }
#pragma unroll
            for (int r = 0; r < kRescanRows; ++r) {
                if (done) break;
                const uint64_t pos = r0 + 64 * r + lane;
                bool hit = false;
                Real tm = 0, tc = 0;
                uint32_t occ = 0;
                uint64_t addr = 0;
                if (cs[r] >= 0) {
                    addr = pos + s_delta[wave][cs[r]];
                    const uint64_t sg = w[r] & kSigBits;
                    const uint64_t f = bfold(q ^ sg);
                    const int dist = __popcll(f);
                    uint32_t mp[4];
                    mismatches(f, mp);
                    hit = dist <= 4 && first_seen(mp, cs[r], s_posmask);
                    if (hit) {
                        occ = static_cast<uint32_t>(w[r] >> 40);
                        if (occ == kOccEscape) occ = static_cast<uint32_t>(a.post[addr] >> 32);
                        const Real ro = to_real<Real>(occ);
                        if (a.sc.calc_mit && dist > 0) tm = mit_closed<Real>(mp, dist, s_om, s_t23) * ro;
                        if (a.sc.calc_cfd) tc = cfd_product<Real>(q, sg, mp, dist, s_cfd) * ro;
                    }
                }
                const uint64_t bal = __ballot(hit);
                if (bal == 0) continue;
                double rm = static_cast<double>(tm), rc = static_cast<double>(tc);
                uint64_t rocc = occ;
                tile_reduce_step<1>(rm, rc, rocc);
                tile_reduce_step<2>(rm, rc, rocc);
                tile_reduce_step<4>(rm, rc, rocc);
                tile_reduce_step<8>(rm, rc, rocc);
                tile_reduce_step<16>(rm, rc, rocc);
                tile_reduce_step<32>(rm, rc, rocc);
                uint32_t last = 63;   // the row's hits up to this lane are taken
                if (!bexit(a.sc.method, dm + rm, dc + rc, a.sc.cut)) {
                    dm += rm;
                    dc += rc;
                    occ_sum += rocc;
                    sites += static_cast<uint64_t>(__popcll(bal));
                } else {
                    uint64_t rest = bal;
                    while (rest) {
                        const uint32_t h = static_cast<uint32_t>(__builtin_ctzll(rest));
                        rest &= rest - 1;
                        dm += static_cast<double>(readlane_real(tm, h));
                        dc += static_cast<double>(readlane_real(tc, h));
                        occ_sum += static_cast<uint32_t>(
                            __builtin_amdgcn_readlane(static_cast<int>(occ), static_cast<int>(h)));
                        sites += 1;
                        if (bexit(a.sc.method, dm, dc, a.sc.cut)) {
                            done = true;
                            last = h;
                            break;
                        }
                    }
                    // (hit by hit the sums may stay below the cut, the row's
                    // butterfly having rounded differently: then the whole
                    // row is taken and the walk goes on)
                }
                if (a.distinct && hit && static_cast<uint32_t>(lane) <= last)
                    atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) +
                                 (static_cast<uint32_t>(a.post[addr]) >> 6),
                             1ULL << (static_cast<uint32_t>(a.post[addr]) & 63));
            }
        }
        if (done) {
            exited = true;
            exit_tile = static_cast<int>(k);
            break;
        }
        // Hit by hit the sums never met the test (the tile's reduced partial
        // rounds differently): the guide goes on with its next tile.
    }
    // Distinct sites of the tiles before the exit tile (or of all tiles).
    if (a.distinct) {
        load_lists();
        const uint64_t P = s_pre[wave][S];
        const uint64_t end = exit_tile >= 0 ? static_cast<uint64_t>(exit_tile) * C : P;
        int cur = 0;
        for (uint64_t b0 = 0; b0 < end; b0 += 64) {
            const uint64_t pos = b0 + lane;
            if (pos >= end) continue;
            while (pos >= s_pre[wave][cur + 1]) ++cur;
            const uint64_t addr = pos + s_delta[wave][cur];
            const uint64_t f = bfold(q ^ (a.site[addr] & kSigBits));
            const int dist = __popcll(f);
            if (dist > 4) continue;
            uint32_t mp[4];
            mismatches(f, mp);
            if (!first_seen(mp, cur, s_posmask)) continue;
            const uint32_t id = static_cast<uint32_t>(a.post[addr]);
            atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (id >> 6), 1ULL << (id & 63));
        }
    }
    if (lane != 0) return;
    if (a.state_out) {
        GuideState st;
        st.m = dm;
        st.c = dc;
        st.occ = occ_sum;
        st.sites = sites | (exited ? kExited : 0);
        a.state_out[g] = st;
        return;
    }
    write_guide(a.sc, a.out, g, dm, dc, occ_sum, sites);
}

// k_finalize without an early exit armed: no rescan can happen, so one lane
// per guide adds the tile partials in tile order, the same additions the
// wave form makes.
constexpr uint32_t kFinSerialTiles = 8;   // more tiles: the guide's wave adds them

// One 64-lane workgroup per 64 guides (blockIdx.x) and gridDim.y replicas of
// it: replica 0 writes the guides with few tiles, one lane each, and replica
// r adds the group's guides with many tiles whose rank among them is r
// modulo the replica count, so that a group of long guides is not added by
// one wave in series.  Each guide is added by one wave in tile order whatever
// the replica count, so the results do not depend on it.
__host__ __device__ inline uint32_t fin_replicas(uint64_t groups) {
    const uint64_t r = 512 / (groups ? groups : 1);
    return static_cast<uint32_t>(r < 1 ? 1 : r > 64 ? 64 : r);
}

__global__ __launch_bounds__(64) void k_finalize_sum(FinArgs a) {
    reset_queue(a.ws);
    const uint64_t g = blockIdx.x * 64ULL + threadIdx.x;
    const int lane = static_cast<int>(threadIdx.x);
    uint32_t K = 0;
    bool mine = false;
    if (g < a.n) {
        K = a.ws.gtiles[g];
        mine = !(K == 0 && a.skip_untiled) && !(K > 1 && a.skip_multi);   // (split: k_ordered_terms)
    }
    // Sums of guide gg over its K tiles, in tile order after its carried state.
    auto finish = [&](uint64_t gg, double dm, double dc, uint64_t occ_sum, uint64_t sites, bool exited) {
        if (a.state_out) {
            GuideState st;
            st.m = dm;
            st.c = dc;
            st.occ = occ_sum;
            st.sites = sites | (exited ? kExited : 0);
            a.state_out[gg] = st;
            return;
        }
        write_guide(a.sc, a.out, gg, dm, dc, occ_sum, sites);
    };
    auto start = [&](uint64_t gg, double &dm, double &dc, uint64_t &occ_sum, uint64_t &sites, bool &exited) {
        dm = 0;
        dc = 0;
        occ_sum = 0;
        sites = 0;
        exited = false;
        if (a.state_in) {
            const GuideState st = a.state_in[gg];
            dm = st.m;
            dc = st.c;
            occ_sum = st.occ;
            sites = st.sites & ~kExited;
            exited = (st.sites & kExited) != 0;
        }
    };
    const uint32_t R = gridDim.y, rep = blockIdx.y;
    // One lane per guide with few tiles.
    if (rep == 0 && mine && K <= kFinSerialTiles) {
        double dm, dc;
        uint64_t occ_sum, sites;
        bool exited;
        start(g, dm, dc, occ_sum, sites, exited);
        for (uint32_t k = 0; k < K; ++k) {
            const TilePart p = a.ws.part[static_cast<uint64_t>(k) * a.n + g];
            dm += p.m;
            dc += p.c;
            occ_sum += p.occ;
            sites += p.sites;
        }
        finish(g, dm, dc, occ_sum, sites, exited);
    }
    // Guides with many tiles: the wave loads 64 partials at a time into LDS,
    // lane 0 adds the MIT sums and lane 1 the CFD sums one by one, in tile
    // order (the same additions, add_chain), and the counts are summed over
    // the lanes (integers: any order).
    __shared__ double2 s_p[64];
    uint32_t rank = 0;
    for (uint64_t m = __ballot(mine && K > kFinSerialTiles); m; m &= m - 1, ++rank) {
        if (rank % R != rep) continue;
        const int l = __builtin_ctzll(m);
        const uint64_t gg = blockIdx.x * 64ULL + static_cast<uint64_t>(l);
        const uint32_t KK = static_cast<uint32_t>(__shfl(static_cast<int>(K), l));
        double dm, dc;
        uint64_t occ_sum, sites;
        bool exited;
        start(gg, dm, dc, occ_sum, sites, exited);
        double ch = lane == 1 ? dc : dm;
        uint64_t po = 0, ps = 0;
        for (uint32_t k0 = 0; k0 < KK; k0 += 64) {
            TilePart p{0.0, 0.0, 0, 0};
            if (k0 + static_cast<uint32_t>(lane) < KK)
                p = a.ws.part[static_cast<uint64_t>(k0 + static_cast<uint32_t>(lane)) * a.n + gg];
            s_p[lane] = make_double2(p.m, p.c);
            po += p.occ;
            ps += p.sites;
            __builtin_amdgcn_wave_barrier();
            if (lane < 2) add_chain<8>(reinterpret_cast<const double *>(s_p) + lane, std::min<uint32_t>(64, KK - k0), ch);
            __builtin_amdgcn_wave_barrier();
        }
        dm = readlane_f64(ch, 0);
        dc = readlane_f64(ch, 1);
        for (int off = 32; off >= 1; off >>= 1) {
            po += shfl64(po, lane ^ off);
            ps += shfl64(ps, lane ^ off);
        }
        occ_sum += po;
        sites += ps;
        if (lane == 0) finish(gg, dm, dc, occ_sum, sites, exited);
    }
}

// In-order split: one wave per split guide adds its hits' terms in the order
// the reference visits them (tile 0's hits, then tile 1's, ...: slices
// ascending, postings in list order) into FP64 sums, one at a time
// (lib/ISSLScoringModule.cpp:365-401; guides are split only with the exit
// off, so every hit counts); it also sums occurrences and sites and marks
// distinct sites.  Lanes fetch
// kOrdTermBatch entries at a time into registers and then LDS, the next
// batch's loads in flight while lanes 0 and 1 add the current one (add_chain).
constexpr int kOrdTermPer = 8;                    // entries per lane per batch
constexpr int kOrdTermBatch = 64 * kOrdTermPer;

__global__ __launch_bounds__(256) void k_ordered_terms(FinArgs a) {
    __shared__ double2 s_t[4][kOrdTermBatch];
    __shared__ uint2 s_o[4][kOrdTermBatch];
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const uint32_t nsplit = a.ws.tctr[kSplitCtr];
    const uint32_t W = gridDim.x * 4;
    for (uint32_t i = blockIdx.x * 4 + static_cast<uint32_t>(wave); i < nsplit; i += W) {
        const uint64_t g = a.ws.slist[i];
        const uint32_t K = a.ws.gtiles[g], C = a.ws.gchunk[g], tb = a.ws.gterm[g];
        // Tile hit counts, lane k holding tile k's (K <= kSplitTiles = 64), so
        // the cursor below moves without loads that would wait behind the
        // batch loads in flight.
        const uint32_t cnt_l = static_cast<uint32_t>(lane) < K
                                   ? static_cast<uint32_t>(a.ws.part[static_cast<uint64_t>(lane) * a.n + g].sites)
                                   : 0u;
        auto count_of = [&](uint32_t kk) {
            return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cnt_l), static_cast<int>(kk)));
        };
        // Cursor: entry j of tile k, whose hits number cnt.
        uint32_t k = 0, j = 0;
        uint32_t cnt = count_of(0);
        while (j >= cnt && k + 1 < K) {
            ++k;
            cnt = count_of(k);
        }
        // (scalar arrays: arrays of vector types are kept in scratch memory)
        double tmv[kOrdTermPer], tcv[kOrdTermPer];
        uint32_t ov[kOrdTermPer], iv[kOrdTermPer];
        uint32_t nb = 0;     // entries of the batch in registers
        bool have = false;   // registers hold a batch
        double m = 0.0, c = 0.0;
        uint64_t occ = 0;
        uint32_t sites = 0;
        for (;;) {
            // The batch in registers goes to LDS; the next one is fetched and
            // its loads stay in flight while lanes 0 and 1 add this one.
            uint32_t ncur = 0;
            if (have) {
#pragma unroll
                for (int r = 0; r < kOrdTermPer; ++r) {
                    const uint32_t e = static_cast<uint32_t>(r * 64 + lane);
                    if (e < nb) {
                        s_t[wave][e] = make_double2(tmv[r], tcv[r]);
                        s_o[wave][e] = make_uint2(ov[r], iv[r]);
                    }
                }
                ncur = nb;
            }
            nb = j < cnt ? std::min<uint32_t>(kOrdTermBatch, cnt - j) : 0u;
            const uint32_t base = tb + k * C + j;
            // Every lane loads without a branch (a load under a branch is
            // waited for at once): entries past the batch re-read its last
            // one, and an empty batch reads the guide's first entry (a split
            // guide has more than 64 entries).
#pragma unroll
            for (int r = 0; r < kOrdTermPer; ++r) {
                const uint32_t e = static_cast<uint32_t>(r * 64 + lane);
                const uint32_t at = nb ? base + std::min(e, nb - 1) : tb;
                const double2 t2 = a.ws.term_mc[at];
                const uint2 o2 = a.ws.term_oi[at];
                tmv[r] = t2.x;
                tcv[r] = t2.y;
                ov[r] = o2.x;
                iv[r] = o2.y;
            }
            j += nb;
            while (j >= cnt && k + 1 < K) {
                ++k;
                j = 0;
                cnt = count_of(k);
            }
            if (!have) {
                have = true;
                if (nb == 0) break;   // no hits at all
                continue;
            }
            __builtin_amdgcn_wave_barrier();
            // Lane 0 adds the MIT terms, lane 1 the CFD terms.
            double ch = lane == 1 ? c : m;
            if (lane < 2) add_chain<16>(reinterpret_cast<const double *>(s_t[wave]) + lane, ncur, ch);
            __builtin_amdgcn_wave_barrier();
            m = readlane_f64(ch, 0);
            c = readlane_f64(ch, 1);
#pragma unroll
            for (int r = 0; r < kOrdTermPer; ++r) {
                const uint32_t e = static_cast<uint32_t>(r * 64 + lane);
                if (e < ncur) {
                    const uint2 v = s_o[wave][e];
                    occ += v.x;
                    sites += 1;
                    if (a.distinct)
                        atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (v.y >> 6),
                                 1ULL << (v.y & 63));
                }
            }
            __builtin_amdgcn_wave_barrier();   // before the next batch overwrites LDS
            if (nb == 0) break;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            occ += shfl64(occ, lane ^ off);
            sites += static_cast<uint32_t>(__shfl(static_cast<int>(sites), lane ^ off));
        }
        if (lane == 0) write_guide(a.sc, a.out, g, m, c, occ, sites);
    }
}

// ------------------------------------------------------------ direct

struct StreamArgs {
    const uint64_t *guides;
    uint64_t n;
    const uint64_t *site;   // signature-inline words
    const uint64_t *post;   // postings (escape occ, ids)
    const uint32_t *lstart;
    const SliceMeta *meta;
    const void *mit;
    const void *cfd;
    int S;
    uint64_t n_sites;
    ScoreScalars sc;
    DevResults out;
    uint64_t *distinct;     // mark scored sites up to the exit (nullable)
    uint64_t heavy;         // guides with more postings go to the tile kernel (0: none)
    uint32_t chunk;         // tile length for those guides
    BatchWs ws;
};

constexpr uint64_t kHanded = ~0ULL;  // plan handed to the tile kernel

// Persistent guide streams.  Wave w scores guides w, w+W, w+2W, ... one after
// another.  For each guide it streams the slice-major site words, DEPTH loads
// per lane in flight, queues the candidates (dist <= 4) in position order in
// an LDS ring and scores them 64 at a time: terms in parallel, then added in
// position order with the exit test after each, as the reference's loop does
// (lib/ISSLScoringModule.cpp:231-401).  In FP64 the sums are therefore the
// reference's, bit for bit.  The next guide is planned while the current one
// streams: its signature is loaded two guides ahead and its list bounds are
// requested right behind the current guide's first round of loads.  Guides
// longer than `heavy` postings leave their plan for the persistent tile
// kernel instead.
template <int PREC, int DEPTH>
__global__ __launch_bounds__(256) void k_stream(StreamArgs a) {
    using Real = typename std::conditional<PREC == ISSL_PREC_FP32C, float, double>::type;
    constexpr int WAVES = 4;
    constexpr uint32_t RING = 64 * (DEPTH + 1);
    __shared__ Real s_cfd[kCfdEntries];
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint32_t s_lbase[kMaxSlices];
    __shared__ uint64_t s_pre[WAVES][2][kMaxSlices + 1];
    __shared__ uint64_t s_delta[WAVES][2][kMaxSlices];
    __shared__ uint32_t s_rpos[WAVES][RING];
    __shared__ uint8_t s_rsl[WAVES][RING];
    __shared__ uint64_t s_posmask[kNoPos + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Real *cfdg = static_cast<const Real *>(a.cfd);
    const Real *mitg = static_cast<const Real *>(a.mit);
    for (int i = tid; i < kCfdEntries; i += 256) s_cfd[i] = cfdg[i];
    if (tid < a.S) {
        s_fold[tid] = a.meta[tid].fold;
        s_lbase[tid] = a.meta[tid].lbase;
    }
    __syncthreads();
    build_posmask(s_posmask, s_fold, a.S, tid);
    __syncthreads();
    const int S = a.S;
    const uint64_t W = gridDim.x * static_cast<uint64_t>(WAVES);
    uint64_t g = blockIdx.x * static_cast<uint64_t>(WAVES) + wave;
    if (g >= a.n) return;  // uniform per wave; no workgroup barrier follows

    // Plan stage 1: list bounds of every slice (lane s) for signature qq.
    auto plan_issue = [&](uint64_t qq, uint32_t &bb, uint32_t &ee) {
        bb = 0;
        ee = 0;
        if (lane < S) {
            const uint32_t key = slice_key(qq, s_fold[lane]);
            const uint32_t lb = s_lbase[lane];
            bb = a.lstart[lb + key];
            ee = a.lstart[lb + key + 1];
        }
    };
    // Plan stage 2: prefix of the list lengths into LDS buffer buf; returns
    // P, or kHanded when the guide goes to the tile kernel.
    auto plan_finish = [&](uint64_t gg, uint32_t bb, uint32_t ee, int buf) -> uint64_t {
        const uint32_t len = ee - bb;
        uint64_t incl = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t v = shfl64(incl, lane >= off ? lane - off : lane);
            if (lane >= off) incl += v;
        }
        const uint64_t excl = incl - len;
        const uint64_t P = S > 0 ? shfl64(incl, S - 1) : 0;
        const uint64_t dl_s = static_cast<uint64_t>(lane) * a.n_sites + bb - excl;
        if (lane < S) {
            s_pre[wave][buf][lane] = excl;
            s_delta[wave][buf][lane] = dl_s;
        }
        if (lane == 0) s_pre[wave][buf][S] = P;  // S may be 64: no lane S
        if (a.heavy) {
            if (P > a.heavy) {
                if (lane < S) {
                    a.ws.pre[gg * (S + 1) + lane] = static_cast<uint32_t>(excl);
                    a.ws.delta[gg * S + lane] = dl_s;
                }
                if (lane == 0) {
                    a.ws.pre[gg * (S + 1) + S] = static_cast<uint32_t>(P);
                    uint64_t c = (P + a.ws.tmax - 1) / a.ws.tmax;
                    c = (c + 63) / 64 * 64;
                    if (c < a.chunk) c = a.chunk;
                    a.ws.gchunk[gg] = static_cast<uint32_t>(c);
                    a.ws.gtiles[gg] = static_cast<uint32_t>((P + c - 1) / c);
                }
                return kHanded;
            }
            if (lane == 0) a.ws.gtiles[gg] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        return P;
    };

    uint64_t qc = a.guides[g] & kSigBits;
    uint64_t Pc;
    {
        uint32_t b0, e0;
        plan_issue(qc, b0, e0);
        Pc = plan_finish(g, b0, e0, 0);
    }
    int cb = 0;
    uint64_t qn = (g + W < a.n) ? (a.guides[g + W] & kSigBits) : 0;

    for (;;) {
        const uint64_t gn = g + W;
        const bool has_next = gn < a.n;
        const uint64_t qnn = (gn + W < a.n) ? (a.guides[gn + W] & kSigBits) : 0;
        uint32_t bn = 0, en = 0;
        bool issued = !has_next;

        if (Pc != kHanded) {
            F2 fm{0.f, 0.f}, fc{0.f, 0.f};
            double dm = 0.0, dc = 0.0;
            uint64_t occ_sum = 0, sites = 0;
            bool done = false;
            uint32_t head = 0, tail = 0;  // wave-uniform ring cursors
            const uint64_t q = qc;

            // Score ring entries [head, head + n): terms per lane, then the
            // running sums in position order with the exit test after each hit.
            auto drain = [&](uint32_t nn) {
                __builtin_amdgcn_wave_barrier();
                bool hit = false;
                Real tm = 0, tc = 0;
                uint32_t occ = 0;
                uint64_t addr = 0;
                if (static_cast<uint32_t>(lane) < nn) {
                    const uint32_t e = (head + lane) % RING;
                    const int sl = s_rsl[wave][e];
                    addr = static_cast<uint64_t>(s_rpos[wave][e]) + s_delta[wave][cb][sl];
                    const uint64_t wv = a.site[addr];
                    const uint64_t sg = wv & kSigBits;
                    const uint64_t f = bfold(q ^ sg);
                    const int dist = __popcll(f);
                    uint32_t mp[4];
                    mismatches(f, mp);
                    hit = first_seen(mp, sl, s_posmask);
                    if (hit) {
                        occ = static_cast<uint32_t>(wv >> 40);
                        if (occ == kOccEscape) occ = static_cast<uint32_t>(a.post[addr] >> 32);
                        const Real ro = to_real<Real>(occ);
                        if (a.sc.calc_mit && dist > 0) tm = mitg[mit_rank(mp, dist)] * ro;
                        if (a.sc.calc_cfd) tc = cfd_product<Real>(q, sg, mp, dist, s_cfd) * ro;
                    }
                }
                uint64_t hb = __ballot(hit);
                uint32_t accepted = nn;  // positions past the exit are not scored
                while (hb) {
                    const int h = __builtin_ctzll(hb);
                    hb &= hb - 1;
                    const Real hm = __shfl(tm, h), hc = __shfl(tc, h);
                    occ_sum += static_cast<uint32_t>(__shfl(static_cast<int>(occ), h));
                    sites += 1;
                    double vm, vc;
                    if constexpr (PREC == ISSL_PREC_FP32C) {
                        fm = f2_add(fm, hm);
                        fc = f2_add(fc, hc);
                        vm = f2_value(fm);
                        vc = f2_value(fc);
                    } else {
                        dm += hm;
                        dc += hc;
                        vm = dm;
                        vc = dc;
                    }
                    if (a.sc.use_cut && bexit(a.sc.method, vm, vc, a.sc.cut)) {
                        done = true;
                        accepted = static_cast<uint32_t>(h) + 1;
                        break;
                    }
                }
                if (a.distinct && hit && static_cast<uint32_t>(lane) < accepted) {
                    const uint32_t id = static_cast<uint32_t>(a.post[addr]);
                    atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (id >> 6),
                             1ULL << (id & 63));
                }
                head += nn;
                __builtin_amdgcn_wave_barrier();
            };

            int cur = 0;
            uint64_t nb = 0, dl = 0;
            if (S > 0) {
                while (cur < S - 1 && static_cast<uint64_t>(lane) >= s_pre[wave][cb][cur + 1]) ++cur;
                nb = s_pre[wave][cb][cur + 1];
                dl = s_delta[wave][cb][cur];
            }
            for (uint64_t base = 0; base < Pc && !done; base += 64 * DEPTH) {
                uint64_t wv[DEPTH];
                int sl[DEPTH];
#pragma unroll
                for (int d = 0; d < DEPTH; ++d) {
                    const uint64_t pos = base + 64 * d + lane;
                    sl[d] = -1;
                    wv[d] = 0;
                    if (pos < Pc) {
                        if (pos >= nb) {
                            do {
                                ++cur;
                                nb = s_pre[wave][cb][cur + 1];
                            } while (pos >= nb);
                            dl = s_delta[wave][cb][cur];
                        }
                        sl[d] = cur;
                        wv[d] = a.site[pos + dl];
                    }
                }
                if (!issued) {  // next guide's list bounds, behind this round's loads
                    plan_issue(qn, bn, en);
                    issued = true;
                }
#pragma unroll
                for (int d = 0; d < DEPTH; ++d) {
                    bool cand = false;
                    if (sl[d] >= 0) cand = __popcll(bfold(q ^ (wv[d] & kSigBits))) <= 4;
                    const uint64_t bb = __ballot(cand);
                    if (bb) {
                        const uint32_t r = __builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(bb >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bb), 0u));
                        if (cand) {
                            const uint32_t e = (tail + r) % RING;
                            s_rpos[wave][e] = static_cast<uint32_t>(base + 64 * d + lane);
                            s_rsl[wave][e] = static_cast<uint8_t>(sl[d]);
                        }
                        tail += static_cast<uint32_t>(__popcll(bb));
                    }
                }
                while (tail - head >= 64 && !done) drain(64);
            }
            while (tail != head && !done) drain(std::min<uint32_t>(64, tail - head));

            if (lane == 0) {
                const double smit = (PREC == ISSL_PREC_FP32C) ? f2_value(fm) : dm;
                const double scfd = (PREC == ISSL_PREC_FP32C) ? f2_value(fc) : dc;
                write_guide(a.sc, a.out, g, smit, scfd, occ_sum, sites);
            }
        }
        if (!has_next) break;
        if (!issued) plan_issue(qn, bn, en);
        Pc = plan_finish(gn, bn, en, cb ^ 1);
        cb ^= 1;
        g = gn;
        qc = qn;
        qn = qnn;
    }
}

// ------------------------------------------------------------ launcher

namespace {

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Sizes of the workspace's arrays for n guides over S slices, with term_cap
// term entries for the in-order split (0: none).
struct WsSizes {
    size_t pre, gres, delta, u32, part, list, ctr, term_mc, term_oi, split, total;
    WsSizes(uint64_t n, int S, bool res, uint32_t tmax, uint64_t term_cap = 0) {
        pre = align256(n * (S + 1) * 4);
        gres = res ? align256(n * std::max(S, 1) * 4) : 0;
        delta = align256(n * std::max(S, 1) * 8);
        u32 = align256(n * 4);
        part = align256(n * tmax * sizeof(TilePart));
        list = align256(kRegions * region_guides(n, S) * (tmax + kLists - 1) * 4);
        ctr = align256(kQueueCtrs * 4);
        term_mc = align256(term_cap * sizeof(double2));
        term_oi = align256(term_cap * sizeof(uint2));
        split = term_cap ? 2 * u32 : 0;   // gterm, slist
        total = ctr + pre + delta + 2 * u32 + part + list + gres + term_mc + term_oi + split;
    }
};

// fresh: the queue counters are not known to be zero (the buffer was
// (re)allocated, or is the caller's).
// A caller's workspace (ws.ext, the allocation-free entry point) keeps the
// smaller tile cap, so that a fixed size holds more guides.
uint32_t ws_tiles_cap(const Workspace &ws, uint64_t n) {
    return ws.ext ? static_cast<uint32_t>(kFewTiles) : tiles_cap(n);
}

int carve(Workspace &ws, uint64_t n, int S, bool res, uint64_t term_cap, BatchWs *b, bool *fresh) {
    const uint32_t tmax = ws_tiles_cap(ws, n);
    const WsSizes z(n, S, res, tmax, term_cap);
    char *p;
    if (ws.ext) {
        if (z.total > ws.ext_bytes)
            return fail(ISSL_ERR_CAPACITY, "the scoring workspace for " + std::to_string(n) +
                                               " guides needs " + std::to_string(z.total) +
                                               " bytes, more than the " +
                                               std::to_string(ws.ext_bytes) + " available");
        p = ws.ext;
        *fresh = true;
    } else {
        const void *old = ws.tiles.p;
        int rc = ws.tiles.reserve(z.total);
        if (rc != ISSL_OK) return rc;
        *fresh = ws.tiles.p != old;
        p = static_cast<char *>(ws.tiles.p);
    }
    // The queue counters come first, at the same place for every n: they are
    // zeroed once and then kept at zero between launch sequences, while the
    // arrays after them move with n.
    b->tctr = reinterpret_cast<uint32_t *>(p);
    p += z.ctr;
    b->pre = reinterpret_cast<uint32_t *>(p);
    p += z.pre;
    b->delta = reinterpret_cast<uint64_t *>(p);
    p += z.delta;
    b->gchunk = reinterpret_cast<uint32_t *>(p);
    p += z.u32;
    b->gtiles = reinterpret_cast<uint32_t *>(p);
    p += z.u32;
    b->part = reinterpret_cast<TilePart *>(p);
    p += z.part;
    b->tlist = reinterpret_cast<uint32_t *>(p);
    p += z.list;
    b->tmax = tmax;
    b->gres = res ? reinterpret_cast<uint32_t *>(p) : nullptr;
    p += z.gres;
    b->term_cap = static_cast<uint32_t>(term_cap);
    b->term_mc = term_cap ? reinterpret_cast<double2 *>(p) : nullptr;
    p += z.term_mc;
    b->term_oi = term_cap ? reinterpret_cast<uint2 *>(p) : nullptr;
    p += z.term_oi;
    b->gterm = term_cap ? reinterpret_cast<uint32_t *>(p) : nullptr;
    p += term_cap ? z.u32 : 0;
    b->slist = term_cap ? reinterpret_cast<uint32_t *>(p) : nullptr;
    return ISSL_OK;
}

// Dynamic LDS of the tile kernel: its S-sized arrays (see k_tiles).
size_t tiles_lds_bytes(int S, bool res) {
    const size_t s = static_cast<size_t>(S);
    size_t b = kTileWaves * 2 * s * 8;                  // delta
    if (res) b += s * kResMaxCodes * 8;                 // rpm
    b += kTileWaves * 2 * (s + 1) * 4;                  // pre
    if (res) b += kTileWaves * 2 * s * 4 + s * kResMaxCodes;   // gres, rabs
    return (b + 15) / 16 * 16;
}

// Workgroups of a tile kernel that one CU holds at once, for its block size
// and dynamic LDS (cached per kernel and LDS size).
int resident_blocks(const void *fn, size_t lds) {
    struct Entry {
        const void *fn;
        size_t lds;
        int blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (const Entry &e : cache)
        if (e.fn == fn && e.lds == lds) return e.blocks;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kTileBlock, lds) != hipSuccess || nb <= 0)
        nb = 1;
    cache.push_back(Entry{fn, lds, nb});
    return nb;
}

// A persistent tile kernel's launch: at most as many workgroups as the CUs
// hold at once (a wave's first tile is fixed by its rank, so every wave must
// be able to start while the others run).
template <typename Kernel>
void launch_resident(Kernel kern, const TileArgs &ta, int grid, int cus, size_t lds, hipStream_t st,
                     hipEvent_t e0, hipEvent_t e1) {
    const int cap = cus * resident_blocks(reinterpret_cast<const void *>(kern), lds);
    hipExtLaunchKernelGGL(kern, dim3(std::max(1, std::min(grid, cap))), dim3(kTileBlock), lds, st, e0,
                          e1, 0, ta);
}

// e0/e1: timing events attached to the dispatch (null when timing is off).
// The RES kernels need the tile queue and the planner's guide residuals.
template <int PREC>
int launch_tiles_depth(const TileArgs &ta, int depth, int grid, int cus, hipStream_t st,
                       hipEvent_t e0, hipEvent_t e1) {
    if (depth != 4 && depth != 8 && depth != 16)
        return fail(ISSL_ERR_ARG, "tuning depth must be 4, 8 or 16");
    const size_t lds_res = tiles_lds_bytes(ta.S, true), lds_inl = tiles_lds_bytes(ta.S, false);
    if (ta.res && ta.ws.tlist && ta.ws.gres) {
        if (depth == 4) launch_resident(k_tiles<PREC, 4, true, true>, ta, grid, cus, lds_res, st, e0, e1);
        if (depth == 8) launch_resident(k_tiles<PREC, 8, true, true>, ta, grid, cus, lds_res, st, e0, e1);
        if (depth == 16) launch_resident(k_tiles<PREC, 16, true, true>, ta, grid, cus, lds_res, st, e0, e1);
    } else if (ta.ws.tlist) {
        if (depth == 4) launch_resident(k_tiles<PREC, 4, true, false>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 8) launch_resident(k_tiles<PREC, 8, true, false>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 16) launch_resident(k_tiles<PREC, 16, true, false>, ta, grid, cus, lds_inl, st, e0, e1);
    } else {
        if (depth == 4) launch_resident(k_tiles<PREC, 4, false, false>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 8) launch_resident(k_tiles<PREC, 8, false, false>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 16) launch_resident(k_tiles<PREC, 16, false, false>, ta, grid, cus, lds_inl, st, e0, e1);
    }
    return ISSL_OK;
}

// The in-order tile kernel (FP64, queued), either site-word layout.
int launch_ordered_depth(const TileArgs &ta, int depth, int grid, int cus, hipStream_t st,
                         hipEvent_t e0, hipEvent_t e1) {
    constexpr int P = ISSL_PREC_FP64;
    if (!ta.ws.tlist) return fail(ISSL_ERR_STATE, "in-order tiles need the tile queue");
    if (depth != 4 && depth != 8 && depth != 16)
        return fail(ISSL_ERR_ARG, "tuning depth must be 4, 8 or 16");
    const size_t lds_res = tiles_lds_bytes(ta.S, true), lds_inl = tiles_lds_bytes(ta.S, false);
    if (ta.res && ta.ws.gres) {
        if (depth == 4) launch_resident(k_tiles<P, 4, true, true, true>, ta, grid, cus, lds_res, st, e0, e1);
        if (depth == 8) launch_resident(k_tiles<P, 8, true, true, true>, ta, grid, cus, lds_res, st, e0, e1);
        if (depth == 16) launch_resident(k_tiles<P, 16, true, true, true>, ta, grid, cus, lds_res, st, e0, e1);
    } else {
        if (depth == 4) launch_resident(k_tiles<P, 4, true, false, true>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 8) launch_resident(k_tiles<P, 8, true, false, true>, ta, grid, cus, lds_inl, st, e0, e1);
        if (depth == 16) launch_resident(k_tiles<P, 16, true, false, true>, ta, grid, cus, lds_inl, st, e0, e1);
    }
    return ISSL_OK;
}

template <int PREC>
int launch_direct_depth(const StreamArgs &da, int depth, uint32_t grid, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1) {
    const dim3 g(grid), b(256);
    switch (depth) {
    case 4: hipExtLaunchKernelGGL((k_stream<PREC, 4>), g, b, 0, st, e0, e1, 0, da); break;
    case 8: hipExtLaunchKernelGGL((k_stream<PREC, 8>), g, b, 0, st, e0, e1, 0, da); break;
    case 16: hipExtLaunchKernelGGL((k_stream<PREC, 16>), g, b, 0, st, e0, e1, 0, da); break;
    default: return fail(ISSL_ERR_ARG, "direct depth must be 4, 8 or 16");
    }
    return ISSL_OK;
}

int cu_count(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (device >= 0 && device < 64) cached[device] = cus;
    return cus;
}

DevResults offset_results(const DevResults &o, uint64_t i) {
    DevResults r = o;
    if (r.mit) r.mit += i;
    if (r.cfd) r.cfd += i;
    if (r.mit_sum) r.mit_sum += i;
    if (r.cfd_sum) r.cfd_sum += i;
    if (r.occ) r.occ += i;
    if (r.sites) r.sites += i;
    if (r.pass) r.pass += i;
    return r;
}

// One launch sequence over n <= kMaxBatch guides.  ordered: the in-order
// form (one tile of length ord_chunk per guide, FP64 sums in visit order,
// the exit taken by the tile kernel); else BATCHED.
int launch_batch(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                 int precision, const issl_tuning &tn, const DevResults &out, uint64_t *distinct,
                 Workspace &ws, hipStream_t st, const Window &win, bool ordered = false,
                 uint32_t ord_chunk = 0, uint32_t split_at = 0, uint64_t term_cap = 0) {
    const int S = static_cast<int>(ix->S);
    const bool carry = win.state_in || win.state_out;
    if (ordered && (carry || precision != ISSL_PREC_FP64))
        return fail(ISSL_ERR_STATE, "the in-order form is FP64 over a resident index");
    // Guide streams (k_stream) read the resident index only.
    const bool direct = !ordered && tn.heavy_kpost != 0 && !carry && win.s0 == 0 && win.s1 == S;
    // The tile kernel streams the list-residual words when the window has
    // them and the planner's queue feeds it.
    const bool res = win.res != nullptr && !direct;
    BatchWs b;
    bool fresh = false;
    const bool split = ordered && split_at && term_cap;
    int rc = carve(ws, n, S, res, split ? term_cap : 0, &b, &fresh);
    if (rc != ISSL_OK) return rc;
    // The counters are zero unless the buffer is new or an earlier sequence
    // stopped before its k_finalize was queued.
    if (fresh || ws.queue_dirty) hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(256), 0, st, b.tctr);
    ws.queue_dirty = true;
    const bool fp32 = precision == ISSL_PREC_FP32C;
    const void *mit = fp32 ? static_cast<const void *>(ix->d_mit32) : static_cast<const void *>(ix->d_mit64);
    const void *cfd = fp32 ? static_cast<const void *>(ix->d_cfd32) : static_cast<const void *>(ix->d_cfd64);
    const bool tiles = !direct || tn.heavy_kpost != kHeavyNever;
    const uint32_t gw = static_cast<uint32_t>((n + 3) / 4);

    StreamArgs da{};
    uint32_t sg = 0;
    auto stream_launch = [&](hipEvent_t e0, hipEvent_t e1) {
        return fp32 ? launch_direct_depth<ISSL_PREC_FP32C>(da, tn.direct_depth, sg, st, e0, e1)
                    : launch_direct_depth<ISSL_PREC_FP64>(da, tn.direct_depth, sg, st, e0, e1);
    };
    if (direct) {
        b.tlist = nullptr;   // guide streams write plans without the queue: slot schedule
        da.guides = d_guides;
        da.n = n;
        da.site = win.inl;   // slice s at s * N: the window is the whole index
        da.post = win.post;
        da.lstart = ix->d_lstart;
        da.meta = ix->d_meta;
        da.mit = mit;
        da.cfd = cfd;
        da.S = S;
        da.n_sites = ix->n;
        da.sc = sc;
        da.out = out;
        da.distinct = distinct;
        da.heavy = tiles ? static_cast<uint64_t>(tn.heavy_kpost) * 1024 : 0;
        da.chunk = static_cast<uint32_t>(tn.chunk);
        da.ws = b;
        // Persistent guide streams: grid_per_cu workgroups per CU, at most one wave per guide.
        const uint64_t sgrid = std::max<uint64_t>(
            1, std::min<uint64_t>(gw, static_cast<uint64_t>(cu_count(ix->device)) * tn.grid_per_cu));
        sg = static_cast<uint32_t>(sgrid);
        if (!tiles) {
            if ((rc = timed_dispatch(ix, st, stream_launch)) != ISSL_OK) return rc;
            ISSL_HIP(hipGetLastError());
            ws.queue_dirty = false;   // no planner ran: the counters are untouched
            return ISSL_OK;
        }
        // (launched below with the tile kernel, the two timed together)
    } else {
        const uint32_t plan_chunk = ordered ? ord_chunk : static_cast<uint32_t>(tn.chunk);
        PlanArgs pa{d_guides, n, ix->d_lstart, ix->d_meta, S, ix->n, plan_chunk,
                    static_cast<uint32_t>(tn.chunk) / 2, win.s0, win.s1, win.state_in, b};
        pa.split_at = split ? split_at : 0;
        pa.split_chunk = static_cast<uint32_t>(tn.chunk);
        const uint64_t G = kPlanWaves * plan_guides_per_wave(S);
        hipLaunchKernelGGL(k_plan, dim3(static_cast<uint32_t>((n + G - 1) / G)),
                           dim3(kPlanBlock), 0, st, pa);
    }

    TileArgs ta;
    ta.guides = d_guides;
    ta.site = win.inl;
    ta.res = res ? win.res : nullptr;
    ta.res_shift = win.res_shift;
    ta.post = win.post;
    ta.meta = ix->d_meta;
    ta.mitf = ix->d_mitf;
    ta.cfd = cfd;
    ta.n = static_cast<uint32_t>(n);
    ta.S = S;
    // (with an exit armed, BATCHED marks in the finalizer, up to the exit)
    ta.mark_distinct = (distinct && (!sc.use_cut || ordered)) ? 1 : 0;
    ta.calc_mit = sc.calc_mit;
    ta.calc_cfd = sc.calc_cfd;
    ta.distinct = distinct;
    ta.sc = sc;
    ta.ws = b;
    // Persistent grid: grid_per_cu workgroups per CU, no more waves than slots.
    uint64_t grid = static_cast<uint64_t>(cu_count(ix->device)) * tn.grid_per_cu;
    grid = std::min<uint64_t>(grid, (n * b.tmax + kTileWaves - 1) / kTileWaves);
    auto tiles_launch = [&](hipEvent_t e0, hipEvent_t e1) {
        const int cus = cu_count(ix->device);
        if (ordered) return launch_ordered_depth(ta, tn.depth, static_cast<int>(grid), cus, st, e0, e1);
        return fp32 ? launch_tiles_depth<ISSL_PREC_FP32C>(ta, tn.depth, static_cast<int>(grid), cus, st, e0, e1)
                    : launch_tiles_depth<ISSL_PREC_FP64>(ta, tn.depth, static_cast<int>(grid), cus, st, e0, e1);
    };
    if (direct) {
        // Long guides go to the tile kernel: one event pair spans both kernels.
        rc = timed_dispatch(ix, st, [&](hipEvent_t e0, hipEvent_t e1) {
            const int r0 = stream_launch(e0, nullptr);
            return r0 != ISSL_OK ? r0 : tiles_launch(nullptr, e1);
        });
    } else {
        rc = timed_dispatch(ix, st, tiles_launch);
    }
    if (rc != ISSL_OK) return rc;

    FinArgs fa;
    fa.guides = d_guides;
    fa.site = win.inl;
    fa.post = win.post;
    fa.meta = ix->d_meta;
    fa.mit = mit;
    fa.cfd = cfd;
    fa.mitf = ix->d_mitf;
    fa.S = S;
    fa.n = n;
    fa.sc = sc;
    fa.out = out;
    fa.distinct = (distinct && sc.use_cut) ? distinct : nullptr;
    fa.skip_untiled = direct ? 1 : 0;
    fa.skip_multi = split ? 1 : 0;
    fa.state_in = win.state_in;
    fa.state_out = win.state_out;
    fa.ws = b;
    if (split) {
        // The split guides, one wave each (their count is on the device).
        FinArgs fo = fa;
        fo.distinct = distinct;
        const uint64_t waves = std::min<uint64_t>(n, static_cast<uint64_t>(cu_count(ix->device)) * 16);
        hipLaunchKernelGGL(k_ordered_terms, dim3(static_cast<uint32_t>((waves + 3) / 4)), dim3(256), 0, st, fo);
    }
    // (in order: the tile kernel took the exit; the finalizer writes each
    // guide from its tile's partial, or from nothing, and sets the queue
    // counters back to zero)
    if (!sc.use_cut || ordered)
        hipLaunchKernelGGL(k_finalize_sum,
                           dim3(static_cast<uint32_t>((n + 63) / 64), fin_replicas((n + 63) / 64)),
                           dim3(64), 0, st, fa);
    else if (fp32)
        hipLaunchKernelGGL((k_finalize<ISSL_PREC_FP32C>), dim3(gw), dim3(256), 0, st, fa);
    else
        hipLaunchKernelGGL((k_finalize<ISSL_PREC_FP64>), dim3(gw), dim3(256), 0, st, fa);
    ISSL_HIP(hipGetLastError());
    ws.queue_dirty = false;   // k_finalize sets the counters back to zero
    return ISSL_OK;
}

}  // namespace

size_t batch_ws_bytes(uint64_t n, int S, bool res) { return WsSizes(n, S, res, kFewTiles).total; }

issl_tuning default_tuning() {
    issl_tuning t{};
    t.chunk = 1024;
    t.block = kTileBlock;
    t.depth = 8;
    t.grid_per_cu = 6;
    t.direct_depth = 8;
    t.heavy_kpost = 0;  // every guide on the tile kernel
    t.ord_depth = 4;
    t.ord_grid = 6;
    t.ord_split = kOrdSplitDefault;
    return t;
}

// Longest guide the in-order form takes in one tile: tile positions are
// kRelBits wide.
constexpr uint64_t kOrdMaxPostings = (1ULL << kRelBits) - 64;

bool ordered_fits(const issl_index *ix) {
    return ix->d_inline && !ix->src && ix->p_max <= kOrdMaxPostings;
}

int launch_ordered(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                   const issl_tuning &tn_in, const DevResults &out, uint64_t *distinct,
                   Workspace &ws, hipStream_t st) {
    if (n == 0) return ISSL_OK;
    if (!ordered_fits(ix)) return fail(ISSL_ERR_STATE, "index too long for the in-order tiles");
    issl_tuning tn = tn_in;
    if (tn.chunk <= 0) tn = default_tuning();
    tn.depth = ix->ord_depth;
    tn.grid_per_cu = tn_in.ovl_grid > 0 ? tn_in.ovl_grid : ix->ord_grid;
    // One tile per guide: the tile covers the longest possible guide.
    const uint32_t ord_chunk =
        static_cast<uint32_t>(std::max<uint64_t>(64, (ix->p_max + 63) / 64 * 64));
    // Guides longer than ord_split are split over tiles of tn.chunk postings
    // when the index can hold such guides, with term entries for at most
    // kOrdTermCap hits per launch (the rest stay one tile).  Not with the exit
    // armed: a split guide streams all its tiles, while one wave stops at the
    // exit, which most long guides reach early.  The allocation-free
    // workspace splits nothing.
    const uint32_t split_at =
        ix->ord_split == kSplitNever ? 0u : static_cast<uint32_t>(ix->ord_split);
    const bool can_split = split_at && !sc.use_cut && !ws.ext && ix->p_max > split_at;
    Window w;
    w.s0 = 0;
    w.s1 = static_cast<int>(ix->S);
    w.post = ix->d_post;
    w.inl = ix->d_inline;
    w.res = ix->res_use ? ix->d_res : nullptr;
    w.res_shift = ix->res_shift;
    for (uint64_t i = 0; i < n; i += kMaxBatch) {
        const uint64_t m = std::min<uint64_t>(kMaxBatch, n - i);
        const uint64_t cap = can_split ? std::min<uint64_t>(kOrdTermCap, m * ix->p_max) : 0;
        const int rc = launch_batch(ix, d_guides + i, m, sc, ISSL_PREC_FP64, tn, offset_results(out, i),
                                    distinct, ws, st, w, true, ord_chunk, split_at, cap);
        if (rc != ISSL_OK) return rc;
    }
    return ISSL_OK;
}

int launch_batched(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                   int precision, int chunk, const issl_tuning &tn_in, const DevResults &out,
                   uint64_t *distinct, Workspace &ws, hipStream_t st, const Window *win) {
    if (n == 0) return ISSL_OK;
    if (!win && !ix->d_inline) return fail(ISSL_ERR_ARG, "this operation needs a resident index");
    if (ix->p_max >= (1ULL << 31))
        return fail(ISSL_ERR_ARG, "index lists are too long for the batched form (a guide could "
                                  "visit 2^31 postings): use the plain form");
    issl_tuning tn = tn_in;
    if (tn.chunk <= 0) tn = default_tuning();
    if (tn_in.ovl_grid > 0) tn.grid_per_cu = tn_in.ovl_grid;   // streams bound to the index
    if (chunk > 0) tn.chunk = chunk;
    if (tn.chunk % 64 != 0 || tn.chunk > (1 << 20))
        return fail(ISSL_ERR_ARG, "chunk must be a multiple of 64 up to 2^20");
    Window w;
    if (win) {
        w = *win;
    } else {
        w.s0 = 0;
        w.s1 = static_cast<int>(ix->S);
        w.post = ix->d_post;
        w.inl = ix->d_inline;
        w.res = ix->res_use ? ix->d_res : nullptr;
        w.res_shift = ix->res_shift;
    }
    for (uint64_t i = 0; i < n; i += kMaxBatch) {
        const uint64_t m = std::min<uint64_t>(kMaxBatch, n - i);
        Window wi = w;
        if (wi.state_in) wi.state_in += i;
        if (wi.state_out) wi.state_out += i;
        const int rc = launch_batch(ix, d_guides + i, m, sc, precision, tn, offset_results(out, i),
                                    distinct, ws, st, wi);
        if (rc != ISSL_OK) return rc;
    }
    return ISSL_OK;
}

}  // namespace issl
