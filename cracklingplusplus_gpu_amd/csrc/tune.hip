// tune.hip -- launch-shape autotuner for the BATCHED form.
//
// Candidates: tile length (chunk), postings in flight per lane (depth),
// persistent workgroups per CU, and the one-wave-per-guide kernel with or
// without hand-off of long guides to the tile kernel.  Each is timed with HIP
// events on a sample of the caller's guides against the resident index, at
// the precision the caller will score in; the fastest is kept on the index
// and, when a cache path is given, appended to a text cache keyed by the
// index shape, the sample size, the precision and the device, so later
// processes skip the sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "device.hpp"

using namespace issl;

namespace {

// The pick depends on the index's shape and on the batch: the sample's size
// enters the key by its power of two.
std::string shape_key(const issl_index *ix, uint64_t n_sample, int precision) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, ix->device) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
    }
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    std::ostringstream k;
    // (p_max, the bound on a guide's postings, tells apart indexes of one
    // site count and slice shape whose lists differ: the in-order split is
    // only tried where guides can pass its threshold)
    k << "t9/" << arch << "/cu" << cus << "/n" << ix->n << "/s" << ix->S << "/pm" << ix->p_max << "/w";
    for (const SliceMeta &m : ix->meta_h) k << m.width << ',';
    int q = 0;
    while ((n_sample >> q) > 1) ++q;
    k << "/q" << q << "/p" << precision;
    return k.str();
}

// A cache line: key, the tuning's fields, the site-word layout (1:
// list-residual words), then the ORDERED form's depth, workgroups per CU and
// split threshold.
bool cache_lookup(const char *path, const std::string &key, issl_tuning *t, int *res) {
    if (!path) return false;
    std::ifstream in(path);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string k;
        issl_tuning v{};
        int r = 0;
        if (ls >> k >> v.chunk >> v.block >> v.depth >> v.grid_per_cu >> v.direct_depth >>
                v.heavy_kpost >> v.us_per_guide >> r >> v.ord_depth >> v.ord_grid >> v.ord_split &&
            k == key) {
            *t = v;
            *res = r;
            found = true;  // the last entry for a key wins
        }
    }
    return found;
}

// Cache eviction between timed trials.  A batch's lists (about 117 MB for
// 10,000 guides on a 100 M-site index) fit in the 256 MB Infinity Cache, so a
// launch repeated on the same sample finds them there, while a caller scoring
// new batches reads them from HBM.  Reading kEvictBytes of the index's own
// site words (read-only; the XOR is stored only if it equals `key`, which
// keeps the reads) pushes the sample's lines out before each timed launch.
constexpr uint64_t kEvictBytes = 512ull << 20;

__global__ __launch_bounds__(256) void k_evict(const uint4 *p, uint64_t n16, uint32_t key,
                                               uint32_t *sink) {
    uint32_t acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) {
        const uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == key) sink[0] = acc;
}

// The largest resident array of site data, up to kEvictBytes of it (nothing
// for a streamed index, whose postings live in host memory).
int evict_caches(const issl_index *ix, uint32_t *sink, hipStream_t st) {
    const void *base = ix->d_inline ? static_cast<const void *>(ix->d_inline)
                                    : static_cast<const void *>(ix->d_post);
    const uint64_t bytes = static_cast<uint64_t>(ix->S) * ix->n * 8;
    if (!base || bytes < 16) return ISSL_OK;
    const uint64_t n16 = std::min<uint64_t>(bytes, kEvictBytes) / 16;
    int cus = 0;
    ISSL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ix->device));
    const uint32_t grid = static_cast<uint32_t>(std::max(1, cus) * 8);
    hipLaunchKernelGGL(k_evict, dim3(grid), dim3(256), 0, st, static_cast<const uint4 *>(base), n16,
                       0x9E3779B9u, sink);
    ISSL_HIP(hipGetLastError());
    return ISSL_OK;
}

}  // namespace

extern "C" int issl_get_tuning(issl_index *ix, issl_tuning *out) {
    if (!ix || !out) return fail(ISSL_ERR_ARG, "null argument");
    *out = ix->tuning.chunk > 0 ? ix->tuning : default_tuning();
    out->ord_depth = ix->ord_depth;
    out->ord_grid = ix->ord_grid;
    out->ord_split = ix->ord_split;
    out->ovl_grid = ix->ovl_grid;
    return ok();
}

extern "C" int issl_set_tuning(issl_index *ix, const issl_tuning *in) {
    if (!ix || !in) return fail(ISSL_ERR_ARG, "null argument");
    const int d = in->depth;
    if (in->chunk < 64 || in->chunk % 64 != 0 || in->chunk > (1 << 20) ||
        !(d == 4 || d == 8 || d == 16) || in->grid_per_cu < 1 || in->grid_per_cu > 32 ||
        !(in->direct_depth == 4 || in->direct_depth == 8 || in->direct_depth == 16) ||
        in->heavy_kpost < 0)
        return fail(ISSL_ERR_ARG, "tuning out of range");
    const int od = in->ord_depth, og = in->ord_grid, os = in->ord_split;
    if (!(od == 0 || od == 4 || od == 8 || od == 16) || og < 0 || og > 32 ||
        !(os == 0 || os >= 64))
        return fail(ISSL_ERR_ARG, "ORDERED tuning out of range");
    if (in->ovl_grid < 0 || in->ovl_grid > 32) return fail(ISSL_ERR_ARG, "ovl_grid out of range");
    std::lock_guard<std::mutex> lk(ix->mu);
    ix->tuning = *in;
    ix->tuning.block = 256;
    ix->tuned = true;
    const issl_tuning def = default_tuning();
    ix->ord_depth = od ? od : def.ord_depth;
    ix->ord_grid = og ? og : def.ord_grid;
    ix->ord_split = os ? os : def.ord_split;
    ix->ovl_grid = in->ovl_grid;
    ix->tuning.ovl_grid = 0;
    return ok();
}

extern "C" int issl_autotune_precision(issl_index *ix, const uint64_t *sample, uint64_t n,
                                       int32_t precision, const char *cache_path,
                                       issl_tuning *best) {
    if (!ix || !best) return fail(ISSL_ERR_ARG, "null argument");
    if (n == 0 || !sample) return fail(ISSL_ERR_ARG, "autotune needs sample guides");
    if (ix->src) return fail(ISSL_ERR_ARG, "autotune needs a resident index");
    if (precision != ISSL_PREC_FP64 && precision != ISSL_PREC_FP32C)
        return fail(ISSL_ERR_ARG, "unknown precision");
    const std::string key = shape_key(ix, n, precision);
    issl_tuning cached{};
    int cached_res = 0;
    if (cache_lookup(cache_path, key, &cached, &cached_res)) {
        cached.ovl_grid = ix->ovl_grid;   // not part of the timed tuning: kept
        int rc = issl_set_tuning(ix, &cached);
        if (rc == ISSL_OK) {
            std::lock_guard<std::mutex> lk(ix->mu);
            ix->res_use = cached_res != 0;
            *best = ix->tuning;
            best->ord_depth = ix->ord_depth;
            best->ord_grid = ix->ord_grid;
            best->ord_split = ix->ord_split;
            return ok();
        }
    }
    std::lock_guard<std::mutex> lk(ix->mu);
    ISSL_HIP(hipSetDevice(ix->device));
    int rc;
    if ((rc = ix->ws.guides.reserve(n * 8)) != ISSL_OK) return rc;
    if ((rc = ix->ws.out.reserve(n * 8 * 2 + 4096)) != ISSL_OK) return rc;
    hipStream_t st = ix->stream;
    if ((rc = ws_acquire(ix, st)) != ISSL_OK) return rc;
    ISSL_HIP(hipMemcpyAsync(ix->ws.guides.p, sample, n * 8, hipMemcpyHostToDevice, st));
    const uint64_t *dg = static_cast<const uint64_t *>(ix->ws.guides.p);
    DevResults o;
    o.mit = static_cast<double *>(ix->ws.out.p);
    o.cfd = o.mit + n;
    ScoreScalars sc{};
    sc.method = ISSL_METHOD_MIT_AND_CFD;
    sc.calc_mit = 1;
    sc.calc_cfd = 1;
    sc.use_cut = 0;
    sc.threshold = 0;
    hipEvent_t e0, e1;
    ISSL_HIP(hipEventCreate(&e0));
    ISSL_HIP(hipEventCreate(&e1));
    // Time one launch shape: two warm-ups, then the best of seven trials of
    // one launch each, with the caches evicted before every trial (the
    // eviction runs before the first event), so each trial reads its lists
    // from HBM as a new batch does.
    constexpr int kTrials = 7;
    uint32_t *const sink = reinterpret_cast<uint32_t *>(o.cfd + n);
    auto measure = [&](issl_tuning &t) -> int {
        int r0;
        for (int r = 0; r < 2; ++r)
            if ((r0 = launch_batched(ix, dg, n, sc, precision, 0, t, o, nullptr, ix->ws, st)) != ISSL_OK)
                return r0;
        float bestms = 1e30f;
        for (int r = 0; r < kTrials; ++r) {
            if ((r0 = evict_caches(ix, sink, st)) != ISSL_OK) return r0;
            ISSL_HIP(hipEventRecord(e0, st));
            if ((r0 = launch_batched(ix, dg, n, sc, precision, 0, t, o, nullptr, ix->ws, st)) != ISSL_OK)
                return r0;
            ISSL_HIP(hipEventRecord(e1, st));
            ISSL_HIP(hipEventSynchronize(e1));
            float ms = 0;
            ISSL_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < bestms) bestms = ms;
        }
        t.us_per_guide = 1000.0 * bestms / static_cast<double>(n);
        return ISSL_OK;
    };
    issl_tuning top{};
    top.us_per_guide = 1e300;
    auto consider = [&](issl_tuning t) -> int {
        const int r0 = measure(t);
        if (r0 != ISSL_OK) return r0;
        if (t.us_per_guide < top.us_per_guide) top = t;
        return ISSL_OK;
    };
    // (1) every guide on the persistent tile kernel: tile shape and occupancy,
    // for each site-word layout the index holds (the list-residual words move
    // half the bytes; they pay where the streams are long enough to be
    // bandwidth-bound, and the signature-inline words where a shorter kernel
    // is bound by latency and issue)
    const int chunks[] = {1024, 2048, 4096};
    const int depths[] = {4, 8, 16};
    const int grids[] = {4, 6, 8};
    // Where the index holds the list-residual words they are used unless the
    // signature-inline words beat them by more than kResMargin, above the
    // trial-to-trial spread of one to two percent: they move half the bytes,
    // which pays wherever the streams are long, and where a launch is bound by
    // latency and issue instead the two layouts time within that spread, so a
    // pick without the margin would follow the noise.
    constexpr double kResMargin = 0.97;
    issl_tuning layout_best[2];
    layout_best[0].us_per_guide = layout_best[1].us_per_guide = 1e300;
    for (int layout = ix->d_res ? 1 : 0; layout >= 0; --layout) {
        ix->res_use = layout != 0;
        for (int c : chunks)
            for (int d : depths)
                for (int gpc : grids) {
                    issl_tuning t{c, 256, d, gpc, 8, 0, 0};
                    if ((rc = measure(t)) != ISSL_OK) return rc;
                    if (t.us_per_guide < layout_best[layout].us_per_guide) layout_best[layout] = t;
                }
    }
    ix->res_use = ix->d_res != nullptr &&
                  !(layout_best[0].us_per_guide < kResMargin * layout_best[1].us_per_guide);
    const issl_tuning tiles_best = layout_best[ix->res_use ? 1 : 0];
    if (tiles_best.us_per_guide < top.us_per_guide) top = tiles_best;
    // (2) one wave per guide, optionally sending long guides to the tile kernel
    const int ddepths[] = {4, 8, 16};
    const int heavies[] = {kHeavyNever, 64, 16, 4};
    const int sgrids[] = {2, 4, 7};
    for (int h : heavies)
        for (int dd : ddepths)
            for (int gpc : sgrids) {
                issl_tuning t = tiles_best;
                t.direct_depth = dd;
                t.heavy_kpost = h;
                t.grid_per_cu = gpc;
                if ((rc = consider(t)) != ISSL_OK) return rc;
            }
    ix->tuning = top;
    ix->tuned = true;
    // (3) the ORDERED form's tile kernel (one wave per guide, hits added in
    // order): load depth and occupancy, on the chosen layout, then the split
    // threshold for long guides when the index can hold guides that long
    if (precision == ISSL_PREC_FP64 && ordered_fits(ix)) {
        auto time_ordered = [&](float *out_ms) -> int {
            float ms_best = 1e30f;
            for (int r = 0; r < 2 + kTrials; ++r) {
                if (r >= 2) {
                    const int r0 = evict_caches(ix, sink, st);
                    if (r0 != ISSL_OK) return r0;
                }
                ISSL_HIP(hipEventRecord(e0, st));
                {
                    const int r0 = launch_ordered(ix, dg, n, sc, top, o, nullptr, ix->ws, st);
                    if (r0 != ISSL_OK) return r0;
                }
                ISSL_HIP(hipEventRecord(e1, st));
                ISSL_HIP(hipEventSynchronize(e1));
                float ms = 0;
                ISSL_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r >= 2 && ms < ms_best) ms_best = ms;   // two warm-up trials
            }
            *out_ms = ms_best;
            return ISSL_OK;
        };
        float best_ms = 1e30f;
        int bd = ix->ord_depth, bg = ix->ord_grid;
        for (int d : depths)
            for (int gpc : grids) {
                ix->ord_depth = d;
                ix->ord_grid = gpc;
                float ms = 0;
                if ((rc = time_ordered(&ms)) != ISSL_OK) return rc;
                if (ms < best_ms) {
                    best_ms = ms;
                    bd = d;
                    bg = gpc;
                }
            }
        ix->ord_depth = bd;
        ix->ord_grid = bg;
        const int splits[] = {kOrdSplitDefault, 16384, kSplitNever};
        int bs = ix->ord_split;
        best_ms = 1e30f;
        for (int sp : splits) {
            if (sp != kOrdSplitDefault && ix->p_max <= 16384) continue;   // no guide that long
            ix->ord_split = sp;
            float ms = 0;
            if ((rc = time_ordered(&ms)) != ISSL_OK) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                bs = sp;
            }
        }
        ix->ord_split = bs;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *best = top;
    best->ord_depth = ix->ord_depth;
    best->ord_grid = ix->ord_grid;
    best->ord_split = ix->ord_split;
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%s %d %d %d %d %d %d %.6g %d %d %d %d\n", key.c_str(), top.chunk, top.block,
                    top.depth, top.grid_per_cu, top.direct_depth, top.heavy_kpost,
                    top.us_per_guide, ix->res_use ? 1 : 0, ix->ord_depth, ix->ord_grid,
                    ix->ord_split);
            fclose(f);
        }
    }
    return ok();
}

extern "C" int issl_autotune(issl_index *ix, const uint64_t *sample, uint64_t n,
                             const char *cache_path, issl_tuning *best) {
    return issl_autotune_precision(ix, sample, n, ISSL_PREC_FP64, cache_path, best);
}
