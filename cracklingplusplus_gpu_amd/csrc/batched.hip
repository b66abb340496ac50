// batched.hip -- BATCHED scoring form: a persistent, work-stealing kernel over
// (guide, chunk) tiles of each guide's slice-major candidate sequence.
//
// Pipeline (all on one stream, capturable in a hipGraph; no memsets):
//   k_plan     one thread per guide: slice keys -> list ranges, prefix P_g,
//              tile length C_g = max(chunk, ceil(P_g/kMaxTiles)) and tiles K_g
//   k_scan     one workgroup: exclusive scan of K_g -> tile_base, total T,
//              queue counter reset
//   k_fillmap  tile -> (guide, k) map
//   k_tiles    persistent waves pop tiles from the atomic queue; each streams
//              its tile's site words (sig | occ<<40, no signature gather),
//              folds, filters, dedups by the slice-order rule and sums its
//              hits; one partial per tile, reduced in a fixed lane order
//   k_finalize one wave per guide: adds the tile partials in tile order and,
//              when the early exit is armed, rescans the tile in which the
//              running total first meets the exit test, hit by hit, to stop
//              exactly where the reference's loop stops
//              (lib/ISSLScoringModule.cpp:365-401)
// Results therefore depend only on the guide and the tile length, never on
// which wave ran which tile or on what else was in the batch.
//
// Precision: ISSL_PREC_FP32C forms each site's MIT/CFD product in FP32 and
// carries every sum as an unevaluated pair (hi, lo) of FP32 built with
// error-free TwoSum; FP64 appears only for hi+lo, the 10000/(100+S)
// transform and the threshold tests.  ISSL_PREC_FP64 uses FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>

#include "device.hpp"

namespace issl {

constexpr int kMaxTiles = 64;        // tiles per guide at most
constexpr int kTileBlock = 256;      // threads per persistent workgroup
constexpr int kTileWaves = kTileBlock / 64;

struct TilePart {
    union {
        struct {
            float mh, ml, ch, cl;
        } f;
        struct {
            double m, c;
        } d;
    };
    uint64_t occ;
    uint64_t sites;
};

// Queue counters sit 128 bytes apart so concurrent atomics on different
// counters never share a line.
constexpr uint32_t kCounterStride = 32;
constexpr uint32_t kMaxWaves = 256 * 32 * kTileWaves;  // CUs x grid/CU x waves

struct BatchWs {
    uint64_t *pre;       // [n][S+1]
    uint64_t *delta;     // [n][S]   posting index = position + delta
    uint32_t *gchunk;    // [n]
    uint32_t *gtiles;    // [n]
    uint32_t *tile_base; // [n+1]
    uint32_t *ctrl;      // [0] global queue counter, [1] total tiles
    uint32_t *cnt;       // [waves * kCounterStride] per-wave range counters
    uint64_t *tile_map;  // [n*kMaxTiles] guide<<32 | k
    TilePart *part;      // [n*kMaxTiles]
};

// Tile scheduling of k_tiles.
//   kSchedQueue : every wave pops the next tile from one global counter
//   kSchedStatic: wave w takes tiles w, w+W, w+2W, ... (no atomics)
//   kSchedSteal : wave w owns tiles [wT/W, (w+1)T/W) behind its own counter,
//                 then steals from up to `steal` other waves' ranges; each
//                 owner drains its own range, so every tile is processed
enum { kSchedQueue = 0, kSchedStatic = 1, kSchedSteal = 2 };

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

__device__ __forceinline__ uint32_t brank(uint64_t f, int k) {
    uint32_t r = (k == 1) ? 0u : (k == 2) ? 20u : (k == 3) ? 210u : 1350u;
    for (int i = 1; i <= k; ++i) {
        const uint32_t p = static_cast<uint32_t>(__builtin_ctzll(f)) >> 1;
        f &= f - 1;
        const uint32_t c = (i == 1)   ? p
                           : (i == 2) ? p * (p - 1) / 2
                           : (i == 3) ? p * (p - 1) * (p - 2) / 6
                                      : p * (p - 1) * (p - 2) * (p - 3) / 24;
        r += c;
    }
    return r;
}

template <typename T>
__device__ __forceinline__ T cfd_product(uint64_t q, uint64_t site, uint64_t f, int dist,
                                         const T *cfd) {
    if (dist == 0) return T(1);
    T c = cfd[320 + 10];
    while (f) {
        const uint32_t p = static_cast<uint32_t>(__builtin_ctzll(f)) >> 1;
        f &= f - 1;
        const uint32_t g = static_cast<uint32_t>(q >> (2 * p)) & 3u;
        const uint32_t o = static_cast<uint32_t>(site >> (2 * p)) & 3u;
        c *= cfd[(p << 4) | (g << 2) | (o ^ 3u)];
    }
    return c;
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
    BatchWs ws;
};

// One wave per guide, one lane per slice: the slice keys, list ranges and
// their prefix are formed in parallel (two dependent loads per guide).
__global__ __launch_bounds__(256) void k_plan(PlanArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t g = blockIdx.x * 4ULL + (threadIdx.x >> 6);
    if (g >= a.n) return;  // uniform per wave
    const uint64_t q = a.guides[g] & kSigBits;
    uint32_t b = 0, len = 0;
    if (lane < a.S) {
        const uint32_t lb = a.meta[lane].lbase;
        const uint32_t key = slice_key(q, a.meta[lane].fold);
        b = a.lstart[lb + key];
        len = a.lstart[lb + key + 1] - b;
    }
    uint64_t incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t v = shfl64(incl, lane >= off ? lane - off : lane);
        if (lane >= off) incl += v;
    }
    const uint64_t excl = incl - len;
    const uint64_t run = a.S > 0 ? shfl64(incl, a.S - 1) : 0;
    if (lane < a.S) {
        a.ws.pre[g * (a.S + 1) + lane] = excl;
        // posting index of position p in slice s: s*N + b + (p - pre[s])
        a.ws.delta[g * a.S + lane] = static_cast<uint64_t>(lane) * a.n_sites + b - excl;
    }
    if (lane == 0) {
        a.ws.pre[g * (a.S + 1) + a.S] = run;
        uint64_t c = (run + kMaxTiles - 1) / kMaxTiles;
        c = (c + 63) / 64 * 64;
        if (c < a.chunk) c = a.chunk;
        a.ws.gchunk[g] = static_cast<uint32_t>(c);
        a.ws.gtiles[g] = static_cast<uint32_t>((run + c - 1) / c);
    }
}

constexpr int kScanBlock = 1024;

// One workgroup: exclusive scan of the per-guide tile counts (tile_base,
// total), the tile -> (guide, k) map, and reset of the per-wave range
// counters (n_counters > 0 only when k_tiles steals).  Each thread owns a
// contiguous run of guides; the run totals are scanned with wave shuffles
// and one LDS pass over the 16 wave totals.
__global__ __launch_bounds__(kScanBlock) void k_scan(BatchWs ws, uint64_t n, uint32_t n_counters) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (uint32_t i = t; i < n_counters; i += kScanBlock) ws.cnt[i * kCounterStride] = 0;
    const uint64_t per = (n + kScanBlock - 1) / kScanBlock;
    const uint64_t lo = std::min<uint64_t>(n, per * t), hi = std::min<uint64_t>(n, lo + per);
    uint32_t sum = 0;
    for (uint64_t i = lo; i < hi; ++i) sum += ws.gtiles[i];
    uint32_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(incl), lane >= off ? lane - off : lane));
        if (lane >= off) incl += v;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    if (wv == 0) {
        uint32_t x = lane < kScanBlock / 64 ? s_wave[lane] : 0u;
#pragma unroll
        for (int off = 1; off < kScanBlock / 64; off <<= 1) {
            const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(x), lane >= off ? lane - off : lane));
            if (lane >= off) x += v;
        }
        if (lane < kScanBlock / 64) s_wave[lane] = x;  // inclusive wave totals
    }
    __syncthreads();
    uint32_t run = (wv > 0 ? s_wave[wv - 1] : 0u) + incl - sum;
    for (uint64_t g = lo; g < hi; ++g) {
        const uint32_t k = ws.gtiles[g];
        ws.tile_base[g] = run;
        for (uint32_t i = 0; i < k; ++i) ws.tile_map[run + i] = (g << 32) | i;
        run += k;
    }
    if (t == kScanBlock - 1) {
        const uint32_t total = s_wave[kScanBlock / 64 - 1];
        ws.tile_base[n] = total;
        ws.ctrl[0] = 0;      // global queue counter
        ws.ctrl[1] = total;  // total tiles
    }
}

// ------------------------------------------------------------ tiles

struct TileArgs {
    const uint64_t *guides;
    const uint64_t *site;   // signature-inline words
    const uint64_t *post;   // postings (escape occ, ids)
    const SliceMeta *meta;
    const void *mit;        // float or double by precision
    const void *cfd;
    int S;
    int mark_distinct;      // mark distinct sites here (no early exit armed)
    int calc_mit, calc_cfd;
    int sched;              // kSched*
    int steal;              // victims tried after the own range (kSchedSteal)
    uint64_t *distinct;
    BatchWs ws;
};

// Postings in flight per lane.
template <int PREC, int DEPTH>
__global__ __launch_bounds__(kTileBlock) void k_tiles(TileArgs a) {
    using Real = typename std::conditional<PREC == ISSL_PREC_FP32C, float, double>::type;
    __shared__ Real s_cfd[336];
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint64_t s_pre[kTileWaves][kMaxSlices + 1];
    __shared__ uint64_t s_delta[kTileWaves][kMaxSlices];
    // Candidate ring: holds < 64 pending plus one round's 64 * DEPTH.
    constexpr uint32_t kRingD = 64 * (DEPTH + 1);
    __shared__ uint32_t s_ring[kTileWaves][kRingD];   // tile-relative position | slice << 26
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Real *cfdg = static_cast<const Real *>(a.cfd);
    const Real *mitg = static_cast<const Real *>(a.mit);
    for (int i = tid; i < 336; i += kTileBlock) s_cfd[i] = cfdg[i];
    if (tid < a.S) s_fold[tid] = a.meta[tid].fold;
    __syncthreads();
    const int S = a.S;
    const uint32_t total = a.ws.ctrl[1];
    const uint32_t W = gridDim.x * kTileWaves;
    const uint32_t w = blockIdx.x * kTileWaves + wave;
    uint32_t iter = 0;   // kSchedStatic
    int tries = 0;       // kSchedSteal: ranges visited after the own one
    uint32_t v = w;      // kSchedSteal: range being drained

    for (;;) {
        uint32_t t = total;
        if (a.sched == kSchedStatic) {
            const uint64_t tt = w + static_cast<uint64_t>(iter++) * W;
            t = tt < total ? static_cast<uint32_t>(tt) : total;
        } else if (a.sched == kSchedSteal) {
            for (;;) {
                const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(v) * total / W);
                const uint32_t hi = static_cast<uint32_t>((static_cast<uint64_t>(v) + 1) * total / W);
                uint32_t k = 0xFFFFFFFFu;
                if (lane == 0 && lo < hi) k = atomicAdd(a.ws.cnt + static_cast<size_t>(v) * kCounterStride, 1u);
                k = static_cast<uint32_t>(__shfl(static_cast<int>(k), 0));
                if (lo < hi && k < hi - lo) {
                    t = lo + k;
                    break;
                }
                if (tries >= a.steal) break;
                ++tries;
                v = static_cast<uint32_t>((w + static_cast<uint64_t>(tries) * 61) % W);
            }
        } else {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(a.ws.ctrl, 1u);
            t = static_cast<uint32_t>(__shfl(static_cast<int>(k), 0));
        }
        if (t >= total) break;
        const uint64_t gm = a.ws.tile_map[t];
        const uint64_t g = gm >> 32;
        const uint32_t k = static_cast<uint32_t>(gm);
        const uint64_t q = a.guides[g] & kSigBits;
        const uint64_t C = a.ws.gchunk[g];
        if (lane < S) s_pre[wave][lane] = a.ws.pre[g * (S + 1) + lane];
        if (lane == 0) s_pre[wave][S] = a.ws.pre[g * (S + 1) + S];  // S may be 64
        if (lane < S) s_delta[wave][lane] = a.ws.delta[g * S + lane];
        __builtin_amdgcn_wave_barrier();
        const uint64_t P = s_pre[wave][S];
        const uint64_t a0 = k * C;
        const uint64_t a1 = std::min<uint64_t>(P, a0 + C);

        Real am = 0, ac = 0;
        F2 fm{0.f, 0.f}, fc{0.f, 0.f};
        uint64_t occ_sum = 0, sites = 0;

        // Candidates (dist <= 4) are queued in a per-wave LDS ring in position
        // order and scored 64 at a time, one per lane, so the rare scoring
        // path runs with every lane busy instead of once per round.  An entry
        // is the tile-relative position and the slice; scoring reloads the
        // site word (a cache hit).
        uint32_t head = 0, tail = 0;  // wave-uniform ring cursors
        auto drain = [&](uint32_t n) {
            __builtin_amdgcn_wave_barrier();
            if (static_cast<uint32_t>(lane) < n) {
                const uint32_t ent = s_ring[wave][(head + lane) % kRingD];
                const int sl = static_cast<int>(ent >> kRelBits);
                const uint64_t addr = a0 + (ent & ((1u << kRelBits) - 1)) + s_delta[wave][sl];
                const uint64_t wv = a.site[addr];
                const uint64_t sg = wv & kSigBits;
                const uint64_t f = bfold(q ^ sg);
                const int dist = __popcll(f);
                bool first = true;
                for (int s2 = 0; s2 < sl; ++s2)
                    if ((f & s_fold[s2]) == 0) {
                        first = false;
                        break;
                    }
                if (first) {
                    uint32_t occ = static_cast<uint32_t>(wv >> 40);
                    if (occ == kOccEscape || a.mark_distinct) {
                        const uint64_t pe = a.post[addr];
                        occ = static_cast<uint32_t>(pe >> 32);
                        if (a.mark_distinct)
                            atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) +
                                         (static_cast<uint32_t>(pe) >> 6),
                                     1ULL << (static_cast<uint32_t>(pe) & 63));
                    }
                    occ_sum += occ;
                    sites += 1;
                    const Real ro = static_cast<Real>(occ);
                    Real tm = 0, tc = 0;
                    if (a.calc_mit && dist > 0) tm = mitg[brank(f, dist)] * ro;
                    if (a.calc_cfd) tc = cfd_product<Real>(q, sg, f, dist, s_cfd) * ro;
                    if constexpr (PREC == ISSL_PREC_FP32C) {
                        fm = f2_add(fm, tm);
                        fc = f2_add(fc, tc);
                    } else {
                        am += tm;
                        ac += tc;
                    }
                }
            }
            head += n;
            __builtin_amdgcn_wave_barrier();
        };

        // Per-lane slice cursor: the next slice boundary and this slice's
        // posting offset stay in registers; LDS is read only at boundaries.
        int cur = 0;
        uint64_t nb, dl;
        {
            const uint64_t p0 = a0 + lane;
            while (cur < S - 1 && p0 >= s_pre[wave][cur + 1]) ++cur;
            nb = s_pre[wave][cur + 1];
            dl = s_delta[wave][cur];
        }
        for (uint64_t base = a0; base < a1; base += 64 * DEPTH) {
            uint64_t wv[DEPTH];
            int sl[DEPTH];
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const uint64_t pos = base + 64 * d + lane;
                sl[d] = -1;
                wv[d] = 0;
                if (pos < a1) {
                    if (pos >= nb) {
                        do {
                            ++cur;
                            nb = s_pre[wave][cur + 1];
                        } while (pos >= nb);
                        dl = s_delta[wave][cur];
                    }
                    sl[d] = cur;
                    wv[d] = a.site[pos + dl];
                }
            }
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                bool cand = false;
                if (sl[d] >= 0) cand = __popcll(bfold(q ^ (wv[d] & kSigBits))) <= 4;
                const uint64_t b = __ballot(cand);
                if (b) {
                    const uint32_t r = __builtin_amdgcn_mbcnt_hi(
                        static_cast<uint32_t>(b >> 32),
                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
                    if (cand) {
                        const uint32_t rel = static_cast<uint32_t>(base - a0) + 64 * d + lane;
                        s_ring[wave][(tail + r) % kRingD] = rel | (static_cast<uint32_t>(sl[d]) << kRelBits);
                    }
                    tail += static_cast<uint32_t>(__popcll(b));
                }
            }
            while (tail - head >= 64) drain(64);
        }
        while (tail != head) drain(std::min<uint32_t>(64, tail - head));
        // Fixed butterfly: every lane ends with the same, order-determined value.
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            occ_sum += shfl64(occ_sum, lane ^ off);
            sites += shfl64(sites, lane ^ off);
            if constexpr (PREC == ISSL_PREC_FP32C) {
                F2 om{__shfl_xor(fm.s, off), __shfl_xor(fm.c, off)};
                F2 oc{__shfl_xor(fc.s, off), __shfl_xor(fc.c, off)};
                fm = f2_merge(fm, om);
                fc = f2_merge(fc, oc);
            } else {
                am += __shfl_xor(am, off);
                ac += __shfl_xor(ac, off);
            }
        }
        if (lane == 0) {
            TilePart tp;
            if constexpr (PREC == ISSL_PREC_FP32C) {
                tp.f.mh = fm.s;
                tp.f.ml = fm.c;
                tp.f.ch = fc.s;
                tp.f.cl = fc.c;
            } else {
                tp.d.m = am;
                tp.d.c = ac;
            }
            tp.occ = occ_sum;
            tp.sites = sites;
            a.ws.part[t] = tp;
        }
        __builtin_amdgcn_wave_barrier();
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
    int S;
    uint64_t n;
    ScoreScalars sc;
    DevResults out;
    uint64_t *distinct;     // mark sites up to the exit (early exit armed)
    int skip_untiled;       // guides without tiles were finished by k_stream
    BatchWs ws;
};

template <int PREC>
__global__ __launch_bounds__(256) void k_finalize(FinArgs a) {
    using Real = typename std::conditional<PREC == ISSL_PREC_FP32C, float, double>::type;
    constexpr int W = 256 / 64;
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint64_t s_pre[W][kMaxSlices + 1];
    __shared__ uint64_t s_delta[W][kMaxSlices];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < a.S) s_fold[tid] = a.meta[tid].fold;
    __syncthreads();
    const uint64_t g = blockIdx.x * static_cast<uint64_t>(W) + wave;
    if (g >= a.n) return;
    const int S = a.S;
    const uint32_t base = a.ws.tile_base[g];
    const uint32_t K = a.ws.tile_base[g + 1] - base;
    if (K == 0 && a.skip_untiled) return;
    TilePart mine{};
    if (lane < static_cast<int>(K)) mine = a.ws.part[base + lane];

    F2 fm{0.f, 0.f}, fc{0.f, 0.f};
    double dm = 0, dc = 0;
    uint64_t occ_sum = 0, sites = 0;
    int exit_tile = -1;
    for (uint32_t k = 0; k < K; ++k) {
        const uint64_t o = shfl64(mine.occ, k);
        const uint64_t si = shfl64(mine.sites, k);
        F2 pm = fm, pc = fc;
        double qm = dm, qc = dc;
        if constexpr (PREC == ISSL_PREC_FP32C) {
            fm = f2_merge(fm, F2{__shfl(mine.f.mh, k), __shfl(mine.f.ml, k)});
            fc = f2_merge(fc, F2{__shfl(mine.f.ch, k), __shfl(mine.f.cl, k)});
        } else {
            dm += __shfl(mine.d.m, k);
            dc += __shfl(mine.d.c, k);
        }
        const double vm = (PREC == ISSL_PREC_FP32C) ? f2_value(fm) : dm;
        const double vc = (PREC == ISSL_PREC_FP32C) ? f2_value(fc) : dc;
        if (a.sc.use_cut && bexit(a.sc.method, vm, vc, a.sc.cut)) {
            // Restore the sums before this tile and rescan it hit by hit.
            fm = pm;
            fc = pc;
            dm = qm;
            dc = qc;
            exit_tile = static_cast<int>(k);
            break;
        }
        occ_sum += o;
        sites += si;
    }

    const uint64_t q = a.guides[g] & kSigBits;
    if (exit_tile >= 0 || a.distinct) {
        if (lane < S) s_pre[wave][lane] = a.ws.pre[g * (S + 1) + lane];
        if (lane == 0) s_pre[wave][S] = a.ws.pre[g * (S + 1) + S];  // S may be 64
        if (lane < S) s_delta[wave][lane] = a.ws.delta[g * S + lane];
        __builtin_amdgcn_wave_barrier();
    }
    const uint64_t C = a.ws.gchunk[g];
    if (exit_tile >= 0) {
        const Real *mitg = static_cast<const Real *>(a.mit);
        const Real *cfdg = static_cast<const Real *>(a.cfd);
        const uint64_t P = s_pre[wave][S];
        const uint64_t a0 = static_cast<uint64_t>(exit_tile) * C;
        const uint64_t a1 = std::min<uint64_t>(P, a0 + C);
        int cur = 0;
        bool done = false;
        for (uint64_t b0 = a0; b0 < a1 && !done; b0 += 64) {
            const uint64_t pos = b0 + lane;
            bool hit = false;
            Real tm = 0, tc = 0;
            uint32_t occ = 0;
            uint64_t addr = 0;
            if (pos < a1) {
                while (pos >= s_pre[wave][cur + 1]) ++cur;
                addr = pos + s_delta[wave][cur];
                const uint64_t w = a.site[addr];
                const uint64_t sg = w & kSigBits;
                const uint64_t f = bfold(q ^ sg);
                const int dist = __popcll(f);
                if (dist <= 4) {
                    hit = true;
                    for (int s2 = 0; s2 < cur; ++s2)
                        if ((f & s_fold[s2]) == 0) {
                            hit = false;
                            break;
                        }
                }
                if (hit) {
                    occ = static_cast<uint32_t>(w >> 40);
                    if (occ == kOccEscape) occ = static_cast<uint32_t>(a.post[addr] >> 32);
                    const Real ro = static_cast<Real>(occ);
                    if (a.sc.calc_mit && dist > 0) tm = mitg[brank(f, dist)] * ro;
                    if (a.sc.calc_cfd) tc = cfd_product<Real>(q, sg, f, dist, cfdg) * ro;
                }
            }
            uint64_t bal = __ballot(hit);
            while (bal) {
                const int h = __builtin_ctzll(bal);
                bal &= bal - 1;
                const Real hm = __shfl(tm, h), hc = __shfl(tc, h);
                const uint32_t ho = static_cast<uint32_t>(__shfl(static_cast<int>(occ), h));
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
                occ_sum += ho;
                sites += 1;
                if (a.distinct && lane == h)
                    atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) +
                                 (static_cast<uint32_t>(a.post[addr]) >> 6),
                             1ULL << (static_cast<uint32_t>(a.post[addr]) & 63));
                if (bexit(a.sc.method, vm, vc, a.sc.cut)) {
                    done = true;
                    break;
                }
            }
        }
    }
    // Distinct sites of the tiles before the exit tile (or all tiles).
    if (a.distinct) {
        const uint64_t P = s_pre[wave][S];
        const uint64_t end = exit_tile >= 0 ? static_cast<uint64_t>(exit_tile) * C : P;
        int cur = 0;
        for (uint64_t b0 = 0; b0 < end; b0 += 64) {
            const uint64_t pos = b0 + lane;
            if (pos >= end) continue;
            while (pos >= s_pre[wave][cur + 1]) ++cur;
            const uint64_t addr = pos + s_delta[wave][cur];
            const uint64_t f = bfold(q ^ (a.site[addr] & kSigBits));
            if (__popcll(f) > 4) continue;
            bool first = true;
            for (int s2 = 0; s2 < cur; ++s2)
                if ((f & s_fold[s2]) == 0) {
                    first = false;
                    break;
                }
            if (!first) continue;
            const uint32_t id = static_cast<uint32_t>(a.post[addr]);
            atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (id >> 6), 1ULL << (id & 63));
        }
    }
    if (lane != 0) return;
    const double smit = (PREC == ISSL_PREC_FP32C) ? f2_value(fm) : dm;
    const double scfd = (PREC == ISSL_PREC_FP32C) ? f2_value(fc) : dc;
    const double vm = 10000.0 / (100.0 + smit);
    const double vc = 10000.0 / (100.0 + scfd);
    const double th = a.sc.threshold;
    bool reject;
    switch (a.sc.method) {
    case ISSL_METHOD_MIT_AND_CFD: reject = vm < th && vc < th; break;
    case ISSL_METHOD_MIT_OR_CFD: reject = vm < th || vc < th; break;
    case ISSL_METHOD_AVG_MIT_CFD: reject = (vm + vc) / 2 < th; break;
    case ISSL_METHOD_MIT: reject = vm < th; break;
    default: reject = vc < th; break;
    }
    const DevResults &o = a.out;
    if (o.mit) o.mit[g] = a.sc.calc_mit ? vm : -1.0;
    if (o.cfd) o.cfd[g] = a.sc.calc_cfd ? vc : -1.0;
    if (o.mit_sum) o.mit_sum[g] = smit;
    if (o.cfd_sum) o.cfd_sum[g] = scfd;
    if (o.occ) o.occ[g] = occ_sum;
    if (o.sites) o.sites[g] = sites;
    if (o.pass) o.pass[g] = reject ? '0' : '1';
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
    __shared__ Real s_cfd[336];
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint32_t s_lbase[kMaxSlices];
    __shared__ uint64_t s_pre[WAVES][2][kMaxSlices + 1];
    __shared__ uint64_t s_delta[WAVES][2][kMaxSlices];
    __shared__ uint32_t s_rpos[WAVES][RING];
    __shared__ uint8_t s_rsl[WAVES][RING];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Real *cfdg = static_cast<const Real *>(a.cfd);
    const Real *mitg = static_cast<const Real *>(a.mit);
    for (int i = tid; i < 336; i += 256) s_cfd[i] = cfdg[i];
    if (tid < a.S) {
        s_fold[tid] = a.meta[tid].fold;
        s_lbase[tid] = a.meta[tid].lbase;
    }
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
                    a.ws.pre[gg * (S + 1) + lane] = excl;
                    a.ws.delta[gg * S + lane] = dl_s;
                }
                if (lane == 0) {
                    a.ws.pre[gg * (S + 1) + S] = P;
                    uint64_t c = (P + kMaxTiles - 1) / kMaxTiles;
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
                    hit = true;
                    for (int s2 = 0; s2 < sl; ++s2)
                        if ((f & s_fold[s2]) == 0) {
                            hit = false;
                            break;
                        }
                    if (hit) {
                        occ = static_cast<uint32_t>(wv >> 40);
                        if (occ == kOccEscape) occ = static_cast<uint32_t>(a.post[addr] >> 32);
                        const Real ro = static_cast<Real>(occ);
                        if (a.sc.calc_mit && dist > 0) tm = mitg[brank(f, dist)] * ro;
                        if (a.sc.calc_cfd) tc = cfd_product<Real>(q, sg, f, dist, s_cfd) * ro;
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
                const double vm = 10000.0 / (100.0 + smit);
                const double vc = 10000.0 / (100.0 + scfd);
                const double th = a.sc.threshold;
                bool reject;
                switch (a.sc.method) {
                case ISSL_METHOD_MIT_AND_CFD: reject = vm < th && vc < th; break;
                case ISSL_METHOD_MIT_OR_CFD: reject = vm < th || vc < th; break;
                case ISSL_METHOD_AVG_MIT_CFD: reject = (vm + vc) / 2 < th; break;
                case ISSL_METHOD_MIT: reject = vm < th; break;
                default: reject = vc < th; break;
                }
                const DevResults &o = a.out;
                if (o.mit) o.mit[g] = a.sc.calc_mit ? vm : -1.0;
                if (o.cfd) o.cfd[g] = a.sc.calc_cfd ? vc : -1.0;
                if (o.mit_sum) o.mit_sum[g] = smit;
                if (o.cfd_sum) o.cfd_sum[g] = scfd;
                if (o.occ) o.occ[g] = occ_sum;
                if (o.sites) o.sites[g] = sites;
                if (o.pass) o.pass[g] = reject ? '0' : '1';
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

int carve(Workspace &ws, uint64_t n, int S, BatchWs *b) {
    const size_t sz_pre = align256(n * (S + 1) * 8);
    const size_t sz_delta = align256(n * std::max(S, 1) * 8);
    const size_t sz_u32 = align256(n * 4);
    const size_t sz_base = align256((n + 1) * 4);
    const size_t sz_ctrl = 256;
    const size_t sz_cnt = align256(static_cast<size_t>(kMaxWaves) * kCounterStride * 4);
    const size_t sz_map = align256(n * kMaxTiles * 8);
    const size_t sz_part = align256(n * kMaxTiles * sizeof(TilePart));
    const size_t total =
        sz_pre + sz_delta + 2 * sz_u32 + sz_base + sz_ctrl + sz_cnt + sz_map + sz_part;
    int rc = ws.tiles.reserve(total);
    if (rc != ISSL_OK) return rc;
    char *p = static_cast<char *>(ws.tiles.p);
    b->pre = reinterpret_cast<uint64_t *>(p);
    p += sz_pre;
    b->delta = reinterpret_cast<uint64_t *>(p);
    p += sz_delta;
    b->gchunk = reinterpret_cast<uint32_t *>(p);
    p += sz_u32;
    b->gtiles = reinterpret_cast<uint32_t *>(p);
    p += sz_u32;
    b->tile_base = reinterpret_cast<uint32_t *>(p);
    p += sz_base;
    b->ctrl = reinterpret_cast<uint32_t *>(p);
    p += sz_ctrl;
    b->cnt = reinterpret_cast<uint32_t *>(p);
    p += sz_cnt;
    b->tile_map = reinterpret_cast<uint64_t *>(p);
    p += sz_map;
    b->part = reinterpret_cast<TilePart *>(p);
    return ISSL_OK;
}

template <int PREC, int DEPTH>
void launch_tiles(const TileArgs &ta, int grid, hipStream_t st) {
    hipLaunchKernelGGL((k_tiles<PREC, DEPTH>), dim3(grid), dim3(kTileBlock), 0, st, ta);
}

template <int PREC>
int launch_tiles_depth(const TileArgs &ta, int depth, int grid, hipStream_t st) {
    switch (depth) {
    case 1: launch_tiles<PREC, 1>(ta, grid, st); break;
    case 2: launch_tiles<PREC, 2>(ta, grid, st); break;
    case 4: launch_tiles<PREC, 4>(ta, grid, st); break;
    case 8: launch_tiles<PREC, 8>(ta, grid, st); break;
    case 16: launch_tiles<PREC, 16>(ta, grid, st); break;
    default: return fail(ISSL_ERR_ARG, "tuning depth must be 1, 2, 4, 8 or 16");
    }
    return ISSL_OK;
}

}  // namespace

issl_tuning default_tuning() {
    issl_tuning t{};
    t.chunk = 2048;
    t.block = kTileBlock;
    t.depth = 8;
    t.grid_per_cu = 4;
    t.sched = kSchedStatic;
    t.steal = 8;
    t.direct_depth = 8;
    t.heavy_kpost = kHeavyNever;
    return t;
}

namespace {

template <int PREC>
int launch_direct_depth(const StreamArgs &da, int depth, uint32_t grid, hipStream_t st) {
    switch (depth) {
    case 4: hipLaunchKernelGGL((k_stream<PREC, 4>), dim3(grid), dim3(256), 0, st, da); break;
    case 8: hipLaunchKernelGGL((k_stream<PREC, 8>), dim3(grid), dim3(256), 0, st, da); break;
    case 16: hipLaunchKernelGGL((k_stream<PREC, 16>), dim3(grid), dim3(256), 0, st, da); break;
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

}  // namespace

int launch_batched(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                   int precision, int chunk, const issl_tuning &tn_in, const DevResults &out,
                   uint64_t *distinct, Workspace &ws, hipStream_t st) {
    if (n == 0) return ISSL_OK;
    if (n > (0xFFFFFFFFULL / kMaxTiles)) return fail(ISSL_ERR_ARG, "too many guides for one launch");
    issl_tuning tn = tn_in;
    if (tn.chunk <= 0) tn = default_tuning();
    if (tn.sched < 0 || tn.sched > kSchedSteal) return fail(ISSL_ERR_ARG, "unknown tile schedule");
    if (chunk > 0) tn.chunk = chunk;
    if (tn.chunk % 64 != 0 || tn.chunk > (1 << 20))
        return fail(ISSL_ERR_ARG, "chunk must be a multiple of 64 up to 2^20");
    const int S = static_cast<int>(ix->S);
    BatchWs b;
    int rc = carve(ws, n, S, &b);
    if (rc != ISSL_OK) return rc;
    const bool fp32 = precision == ISSL_PREC_FP32C;
    const void *mit = fp32 ? static_cast<const void *>(ix->d_mit32) : static_cast<const void *>(ix->d_mit64);
    const void *cfd = fp32 ? static_cast<const void *>(ix->d_cfd32) : static_cast<const void *>(ix->d_cfd64);
    const bool direct = tn.heavy_kpost != 0;
    const bool tiles = !direct || tn.heavy_kpost != kHeavyNever;

    // Persistent grid: grid_per_cu workgroups per CU, no more waves than tiles can need.
    const uint64_t max_tiles = n * kMaxTiles;
    uint64_t grid = static_cast<uint64_t>(cu_count(ix->device)) * tn.grid_per_cu;
    grid = std::min<uint64_t>(grid, kMaxWaves / kTileWaves);
    const uint64_t waves_needed = std::max<uint64_t>(1, std::min<uint64_t>(max_tiles, grid * kTileWaves));
    grid = std::min<uint64_t>(grid, (waves_needed + kTileWaves - 1) / kTileWaves);
    const uint32_t n_waves = static_cast<uint32_t>(grid * kTileWaves);
    const uint32_t gw = static_cast<uint32_t>((n + 3) / 4);

    if (direct) {
        StreamArgs da;
        da.guides = d_guides;
        da.n = n;
        da.site = ix->d_inline;
        da.post = ix->d_post;
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
        timing_mark(ix, st);
        // Persistent guide streams: grid_per_cu workgroups per CU, at most one wave per guide.
        const uint64_t sgrid = std::max<uint64_t>(
            1, std::min<uint64_t>(gw, static_cast<uint64_t>(cu_count(ix->device)) * tn.grid_per_cu));
        rc = fp32 ? launch_direct_depth<ISSL_PREC_FP32C>(da, tn.direct_depth, static_cast<uint32_t>(sgrid), st)
                  : launch_direct_depth<ISSL_PREC_FP64>(da, tn.direct_depth, static_cast<uint32_t>(sgrid), st);
        if (rc != ISSL_OK) return rc;
        timing_mark(ix, st);
        if (!tiles) {
            ISSL_HIP(hipGetLastError());
            return ISSL_OK;
        }
    } else {
        PlanArgs pa{d_guides, n, ix->d_lstart, ix->d_meta, S, ix->n, static_cast<uint32_t>(tn.chunk), b};
        hipLaunchKernelGGL(k_plan, dim3(gw), dim3(256), 0, st, pa);
    }
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanBlock), 0, st, b, n,
                       tn.sched == kSchedSteal ? n_waves : 0u);

    TileArgs ta;
    ta.guides = d_guides;
    ta.site = ix->d_inline;
    ta.post = ix->d_post;
    ta.meta = ix->d_meta;
    ta.mit = mit;
    ta.cfd = cfd;
    ta.S = S;
    ta.mark_distinct = (distinct && !sc.use_cut) ? 1 : 0;
    ta.calc_mit = sc.calc_mit;
    ta.calc_cfd = sc.calc_cfd;
    ta.sched = tn.sched;
    ta.steal = tn.steal;
    ta.distinct = distinct;
    ta.ws = b;
    if (!direct) timing_mark(ix, st);
    rc = fp32 ? launch_tiles_depth<ISSL_PREC_FP32C>(ta, tn.depth, static_cast<int>(grid), st)
              : launch_tiles_depth<ISSL_PREC_FP64>(ta, tn.depth, static_cast<int>(grid), st);
    if (rc != ISSL_OK) return rc;
    if (!direct) timing_mark(ix, st);

    FinArgs fa;
    fa.guides = d_guides;
    fa.site = ix->d_inline;
    fa.post = ix->d_post;
    fa.meta = ix->d_meta;
    fa.mit = mit;
    fa.cfd = cfd;
    fa.S = S;
    fa.n = n;
    fa.sc = sc;
    fa.out = out;
    fa.distinct = (distinct && sc.use_cut) ? distinct : nullptr;
    fa.skip_untiled = direct ? 1 : 0;
    fa.ws = b;
    if (fp32)
        hipLaunchKernelGGL((k_finalize<ISSL_PREC_FP32C>), dim3(gw), dim3(256), 0, st, fa);
    else
        hipLaunchKernelGGL((k_finalize<ISSL_PREC_FP64>), dim3(gw), dim3(256), 0, st, fa);
    ISSL_HIP(hipGetLastError());
    return ISSL_OK;
}

}  // namespace issl
