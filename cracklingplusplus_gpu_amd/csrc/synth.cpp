// synth.cpp -- seeded synthetic workloads: distinct sites with ids in
// sorted-text order (what ExtractOfftargets + sort + ISSLCreateIndex hand the
// scorer), slice masks and query guides.
//
// Every random draw is a function of (seed, stream, item), so outputs do not
// depend on the thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "issl_synth.h"

namespace issl {
namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// splitmix64 stream seeded from (seed, stream, item).
struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream, uint64_t item)
        : s(mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ULL) ^ mix64(item + 0x632BE59BD9B4E019ULL))) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

inline uint64_t prob_to_u64(double p) {
    if (p <= 0) return 0;
    if (p >= 1) return ~0ULL;
    return static_cast<uint64_t>(std::ldexp(p, 64));
}

// Nucleotide draw at a GC fraction: A and T each (1-gc)/2, C and G each gc/2.
struct Base {
    uint64_t ta, tc, tg;
    bool uniform;
    explicit Base(double gc) {
        const double at = (1.0 - gc) / 2, cg = gc / 2;
        ta = prob_to_u64(at);
        tc = prob_to_u64(at + cg);
        tg = prob_to_u64(at + 2 * cg);
        uniform = std::fabs(gc - 0.5) < 1e-12;
    }
    unsigned draw(Rng &r) const {
        const uint64_t x = r.next();
        return x < ta ? 0u : x < tc ? 1u : x < tg ? 2u : 3u;
    }
    uint64_t sig(Rng &r) const {
        if (uniform) return r.next() & ((1ULL << 40) - 1);
        uint64_t s = 0;
        for (unsigned j = 0; j < 20; ++j) s |= static_cast<uint64_t>(draw(r)) << (2 * j);
        return s;
    }
};

// Each position independently replaced by a different base with prob p:
// Binomial(20, p) substitutions at uniformly random positions.
uint64_t substitute(uint64_t sig, uint64_t tp, Rng &r) {
    for (unsigned j = 0; j < 20; ++j)
        if (r.next() < tp) {
            const uint64_t c = (sig >> (2 * j)) & 3ULL;
            const uint64_t nc = (c + 1 + r.below(3)) & 3ULL;
            sig = (sig & ~(3ULL << (2 * j))) | (nc << (2 * j));
        }
    return sig;
}

// Exactly k substitutions at distinct positions.
uint64_t substitute_k(uint64_t sig, int k, Rng &r) {
    int pos[20];
    for (int j = 0; j < 20; ++j) pos[j] = j;
    for (int i = 0; i < k; ++i) {
        const int t = i + static_cast<int>(r.below(20 - i));
        std::swap(pos[i], pos[t]);
        const int j = pos[i];
        const uint64_t c = (sig >> (2 * j)) & 3ULL;
        const uint64_t nc = (c + 1 + r.below(3)) & 3ULL;
        sig = (sig & ~(3ULL << (2 * j))) | (nc << (2 * j));
    }
    return sig;
}

// 1 + Geometric(p), failures convention: P(F = k) = (1-p)^k p.
uint32_t occurrences(double p, Rng &r) {
    const uint64_t tp = prob_to_u64(p);
    uint32_t f = 0;
    while (f < (1u << 20) && r.next() >= tp) ++f;
    return 1 + f;
}

// Position 0 most significant: ascending text key == lexicographic order of
// the 20-mer strings (A < C < G < T), the order ISSLCreateIndex assigns ids in.
inline uint64_t text_key(uint64_t sig) {
    uint64_t k = 0;
    for (unsigned j = 0; j < 20; ++j) k |= ((sig >> (2 * j)) & 3ULL) << (2 * (19 - j));
    return k;
}
inline uint64_t sig_of_text_key(uint64_t k) { return text_key(k); }

// Parallel stable LSD radix sort of 64-bit keys on their low `bits` bits.
void radix_sort(std::vector<uint64_t> &a, int bits, int threads) {
    const int D = 14;
    const uint64_t R = 1ULL << D;
    std::vector<uint64_t> tmp(a.size());
    const uint64_t n = a.size();
    for (int shift = 0; shift < bits; shift += D) {
        std::vector<uint64_t> hist(static_cast<size_t>(threads) * R, 0);
#pragma omp parallel num_threads(threads)
        {
            const int t = omp_get_thread_num();
            const uint64_t lo = n * t / threads, hi = n * (t + 1) / threads;
            uint64_t *h = hist.data() + static_cast<size_t>(t) * R;
            for (uint64_t i = lo; i < hi; ++i) ++h[(a[i] >> shift) & (R - 1)];
#pragma omp barrier
#pragma omp single
            {
                uint64_t run = 0;
                for (uint64_t d = 0; d < R; ++d)
                    for (int tt = 0; tt < threads; ++tt) {
                        uint64_t &c = hist[static_cast<size_t>(tt) * R + d];
                        const uint64_t v = c;
                        c = run;
                        run += v;
                    }
            }
            for (uint64_t i = lo; i < hi; ++i) tmp[h[(a[i] >> shift) & (R - 1)]++] = a[i];
        }
        a.swap(tmp);
    }
}

enum : uint64_t {
    kStreamSeeds = 1,
    kStreamPlanted,
    kStreamRepeat,
    kStreamUnique,
    kStreamOcc,
    kStreamMasks,
    kStreamGuideSite,
    kStreamGuideRepeat,
    kStreamGuideRandom,
    kStreamShuffle,
    kStreamTopUp = 100,
};

}  // namespace
}  // namespace issl

using namespace issl;

extern "C" int issl_synth(const issl_synth_cfg *cfg, uint64_t *sigs, uint32_t *occ,
                          uint64_t *masks, uint64_t *guides, uint8_t *kinds, int threads) {
    if (!cfg || !sigs || !occ) return fail(ISSL_ERR_ARG, "issl_synth: null buffer");
    const uint64_t N = cfg->n_sites;
    const uint64_t Q = cfg->n_guides_site + cfg->n_guides_unique + cfg->n_guides_planted +
                       cfg->n_guides_repeat + cfg->n_guides_random;
    if (Q && !guides) return fail(ISSL_ERR_ARG, "issl_synth: null guides");
    if (cfg->guide_batches < 0 || cfg->guide_batches > 4096)
        return fail(ISSL_ERR_ARG, "issl_synth: guide_batches out of range");
    if (cfg->n_slices && !masks) return fail(ISSL_ERR_ARG, "issl_synth: null masks");
    if (N == 0 || N >= (1ULL << 32)) return fail(ISSL_ERR_ARG, "issl_synth: n_sites out of range");
    if (cfg->n_guides_planted > N) return fail(ISSL_ERR_ARG, "issl_synth: too many planted guides");
    if (cfg->repeat_frac < 0 || cfg->repeat_frac > 1 || cfg->gc < 0 || cfg->gc > 1 ||
        cfg->guide_sub_p < 0 || cfg->guide_sub_p > 1)
        return fail(ISSL_ERR_ARG, "issl_synth: fraction out of range");
    if (cfg->slice_width > 16 || (cfg->contiguous_masks && cfg->n_slices * cfg->slice_width > 20) ||
        cfg->slice_width > 20)
        return fail(ISSL_ERR_ARG, "issl_synth: slice layout does not fit 20 positions");
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    const uint64_t seed = cfg->seed;
    const Base base(cfg->gc);
    const uint64_t tsub = prob_to_u64(cfg->repeat_sub_p);
    const uint64_t tsub_guide = cfg->guide_sub_p > 0 ? prob_to_u64(cfg->guide_sub_p) : tsub;

    // Repeat seeds and planted guides.
    const uint64_t nseeds = cfg->repeat_frac > 0 ? std::max<uint64_t>(1, cfg->n_repeat_seeds) : 0;
    std::vector<uint64_t> rseed(nseeds);
    for (uint64_t i = 0; i < nseeds; ++i) {
        Rng r(seed, kStreamSeeds, i);
        rseed[i] = base.sig(r);
    }
    const uint64_t P = cfg->n_guides_planted;
    std::vector<uint64_t> planted(P), variant(P);
    for (uint64_t i = 0; i < P; ++i) {
        Rng r(seed, kStreamPlanted, i);
        planted[i] = Base(0.5).sig(r);
        variant[i] = substitute_k(planted[i], 1 + static_cast<int>(r.below(4)), r);
    }

    // Pool of text keys, bit 0 = repeat-derived.  Targets: R derived sites,
    // N - R others (planted variants count among the others).
    const uint64_t R = nseeds ? static_cast<uint64_t>(std::llround(cfg->repeat_frac * N)) : 0;
    std::vector<uint64_t> pool;
    pool.reserve(N + N / 64 + 1024);
    for (uint64_t i = 0; i < P; ++i) pool.push_back(text_key(variant[i]) << 1);
    std::sort(pool.begin(), pool.end());
    pool.erase(std::unique(pool.begin(), pool.end()), pool.end());
    uint64_t have_rep = 0, have_uni = pool.size(), rep_drawn = 0, uni_drawn = 0;
    for (int round = 0; round < 64; ++round) {
        // Missing derived sites first; i.i.d. sites fill up to N.  Merging
        // duplicates only shrinks the pool, so it never exceeds N.
        const uint64_t need_rep = R > have_rep ? R - have_rep : 0;
        const uint64_t after = have_rep + have_uni + need_rep;
        const uint64_t need_uni = N > after ? N - after : 0;
        if (need_rep == 0 && need_uni == 0) break;
        const size_t old = pool.size();
        pool.resize(old + need_rep + need_uni);
        const uint64_t rep0 = rep_drawn, uni0 = uni_drawn;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(need_rep); ++i) {
            Rng r(seed, kStreamRepeat, rep0 + i);
            const uint64_t s = substitute(rseed[r.below(nseeds)], tsub, r);
            pool[old + i] = (text_key(s) << 1) | 1ULL;
        }
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(need_uni); ++i) {
            Rng r(seed, kStreamUnique, uni0 + i);
            pool[old + need_rep + i] = text_key(base.sig(r)) << 1;
        }
        rep_drawn += need_rep;
        uni_drawn += need_uni;
        radix_sort(pool, 41, nt);
        // Merge duplicates; a site drawn as a repeat member stays derived.
        size_t w = 0;
        for (size_t i = 0; i < pool.size(); ++i) {
            if (w && (pool[w - 1] >> 1) == (pool[i] >> 1))
                pool[w - 1] |= pool[i] & 1ULL;
            else
                pool[w++] = pool[i];
        }
        pool.resize(w);
        have_rep = 0;
        for (size_t i = 0; i < w; ++i) have_rep += pool[i] & 1ULL;
        have_uni = w - have_rep;
        if (have_rep + have_uni == N && have_rep >= R) break;
    }
    if (pool.size() != N) return fail(ISSL_ERR_STATE, "issl_synth: could not reach n_sites distinct");

#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t id = 0; id < static_cast<int64_t>(N); ++id) {
        const uint64_t tk = pool[id] >> 1;
        sigs[id] = sig_of_text_key(tk);
        Rng r(seed, kStreamOcc, tk);
        occ[id] = occurrences((pool[id] & 1ULL) ? cfg->occ_p_repeat : cfg->occ_p_unique, r);
    }

    // Slice masks.
    for (uint64_t s = 0; s < cfg->n_slices; ++s) {
        uint64_t m = 0;
        if (cfg->contiguous_masks) {
            for (uint64_t j = 0; j < cfg->slice_width; ++j) m |= 1ULL << (s * cfg->slice_width + j);
        } else {
            Rng r(seed, kStreamMasks, s);
            int pos[20];
            for (int j = 0; j < 20; ++j) pos[j] = j;
            for (uint64_t i = 0; i < cfg->slice_width; ++i) {
                const uint64_t t = i + r.below(20 - i);
                std::swap(pos[i], pos[t]);
                m |= 1ULL << pos[i];
            }
        }
        masks[s] = m;
    }

    // Guides: batch b draws items b*count + i of each guide stream and is
    // shuffled on its own, so batches are disjoint draws of the same mix and
    // batch 0 is the single-batch workload.
    const uint64_t B = cfg->guide_batches > 0 ? static_cast<uint64_t>(cfg->guide_batches) : 1;
    std::vector<uint8_t> kind(B * Q);
    uint64_t g = 0;
    for (uint64_t b = 0; b < B; ++b) {
        const uint64_t g0 = g;
        for (uint64_t i = 0; i < cfg->n_guides_site; ++i, ++g) {
            Rng r(seed, kStreamGuideSite, b * cfg->n_guides_site + i);
            guides[g] = sigs[r.below(N)];
            kind[g] = 0;
        }
        for (uint64_t i = 0; i < cfg->n_guides_unique; ++i, ++g) {
            Rng r(seed, kStreamGuideSite, (1ULL << 40) + b * cfg->n_guides_unique + i);
            uint64_t id = r.below(N);
            for (int t = 0; t < 1000 && (pool[id] & 1ULL); ++t) id = r.below(N);
            guides[g] = sigs[id];
            kind[g] = 1;
        }
        for (uint64_t i = 0; i < P; ++i, ++g) {
            guides[g] = planted[i];
            kind[g] = 2;
        }
        for (uint64_t i = 0; i < cfg->n_guides_repeat; ++i, ++g) {
            Rng r(seed, kStreamGuideRepeat, b * cfg->n_guides_repeat + i);
            guides[g] = nseeds ? substitute(rseed[r.below(nseeds)], tsub_guide, r) : base.sig(r);
            kind[g] = 3;
        }
        for (uint64_t i = 0; i < cfg->n_guides_random; ++i, ++g) {
            Rng r(seed, kStreamGuideRandom, b * cfg->n_guides_random + i);
            guides[g] = base.sig(r);
            kind[g] = 4;
        }
        Rng sh(seed, kStreamShuffle, b);
        for (uint64_t i = Q; i > 1; --i) {
            const uint64_t j = sh.below(i);
            std::swap(guides[g0 + i - 1], guides[g0 + j]);
            std::swap(kind[g0 + i - 1], kind[g0 + j]);
        }
    }
    if (kinds) std::memcpy(kinds, kind.data(), B * Q);
    return ok();
}
