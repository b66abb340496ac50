/*
 * issl_oracle.c -- CPU restatement of the reference ISSL scorer.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker or the timed CPU baseline; the
 * product never uses it.  See issl_oracle.h for how each function is pinned.
 *
 * Build: -O2 -fopenmp -ffp-contract=off, no -ffast-math, so the FP64 sums
 * are formed exactly like the reference's (x86-64 baseline has no FMA).
 */
#include "issl_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define FOLD_LO 0x5555555555555555ULL

/* ---------------------------------------------------------------- encode */

/* lib/ISSLScoringModule.cpp:10 maps 'C'->1, 'G'->2, 'T'->3 and every other
 * index below 85 to 0. */
static unsigned nt_code(unsigned char c)
{
    switch (c) {
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default:  return 0;
    }
}

/* lib/ISSLScoringModule.cpp:469-476 (always 20 characters, :219). */
uint64_t or_encode(const char *seq20)
{
    uint64_t sig = 0;
    for (unsigned j = 0; j < 20; ++j)
        sig |= (uint64_t)nt_code((unsigned char)seq20[j]) << (2 * j);
    return sig;
}

/* ------------------------------------------------------------------- MIT */

/* Hsu et al. 2013 position weights, as in ISSLCreateIndex.cpp:239. */
static const double mit_weight[20] = {
    0.0, 0.0, 0.014, 0.0, 0.0, 0.395, 0.317, 0.0, 0.389, 0.079,
    0.445, 0.508, 0.613, 0.851, 0.732, 0.828, 0.615, 0.804, 0.685, 0.583};

/* ISSLCreateIndex.cpp:235-261: the product of (1 - w[p]) over mismatches,
 * times 1/((19 - mean gap)/19*4 + 1), times 1/k^2, times 100, evaluated in
 * the same operation order so the doubles are identical. */
double or_mit_local(const uint64_t *p, uint64_t k)
{
    double t1 = 1.0;
    for (uint64_t i = 0; i < k; ++i)
        t1 = t1 * (1.0 - mit_weight[p[i]]);
    double gap = 0.0;
    if (k == 1) {
        gap = 19.0;
    } else {
        for (uint64_t i = 0; i + 1 < k; ++i)
            gap += (double)(p[i + 1] - p[i]);
        gap = gap / (double)(k - 1);
    }
    double t2 = 1.0 / ((19.0 - gap) / 19.0 * 4.0 + 1);
    double t3 = 1.0 / (double)(k * k);
    return t1 * t2 * t3 * 100;
}

/* ISSLCreateIndex.cpp:271-282 for 20-mers. */
double or_mit_fold(uint64_t fold)
{
    uint64_t pos[20];
    uint64_t k = 0;
    for (uint64_t j = 0; j < 20; ++j)
        if ((fold >> (2 * j)) & 3ULL)
            pos[k++] = j;
    if (k == 0)
        return 0.0;
    return or_mit_local(pos, k);
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Every fold mask with 1..4 of the 20 positions set: 20+190+1140+4845. */
size_t or_mit_table(uint64_t *keys, double *vals)
{
    const size_t total = 6195;
    if (!keys && !vals)
        return total;
    uint64_t *k = (uint64_t *)malloc(total * sizeof(uint64_t));
    size_t n = 0;
    for (uint64_t bits = 1; bits < (1ULL << 20); ++bits) {
        if (__builtin_popcountll(bits) > 4)
            continue;
        uint64_t fold = 0;
        for (unsigned j = 0; j < 20; ++j)
            if ((bits >> j) & 1)
                fold |= 1ULL << (2 * j);
        k[n++] = fold;
    }
    qsort(k, n, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < n; ++i) {
        if (keys)
            keys[i] = k[i];
        if (vals)
            vals[i] = or_mit_fold(k[i]);
    }
    free(k);
    return n;
}

/* ------------------------------------------------------------------- CFD */

/* Doench et al. 2016 CFD penalties, the values of
 * lib/otScorePenalties.cpp:6201-6527.  Index (pos<<4)|(guide_nt<<2)|rc(off_nt),
 * one row per position, four groups of four per guide nucleotide. */
static const double cfd_pos[320] = {
    /* pos  0 */ 1.0000000000, 1.0000000000, 0.8571428570, 1.0000000000,  1.0000000000, 0.9130434780, 1.0000000000, 1.0000000000,  1.0000000000, 1.0000000000, 0.7142857140, 0.9000000000,  1.0000000000, 0.9565217390, 0.8571428570, 1.0000000000,
    /* pos  1 */ 0.7272727270, 0.8000000000, 0.7857142860, 1.0000000000,  0.9090909090, 0.6956521740, 1.0000000000, 0.7272727270,  0.6363636360, 1.0000000000, 0.6923076920, 0.8461538460,  1.0000000000, 0.8400000000, 0.8571428570, 0.8461538460,
    /* pos  2 */ 0.7058823530, 0.6111111110, 0.4285714290, 1.0000000000,  0.6875000000, 0.5000000000, 1.0000000000, 0.8666666670,  0.5000000000, 1.0000000000, 0.3846153850, 0.7500000000,  1.0000000000, 0.5000000000, 0.4285714290, 0.7142857140,
    /* pos  3 */ 0.6363636360, 0.6250000000, 0.3529411760, 1.0000000000,  0.8000000000, 0.5000000000, 1.0000000000, 0.8421052630,  0.3636363640, 1.0000000000, 0.5294117650, 0.9000000000,  1.0000000000, 0.6250000000, 0.6470588240, 0.4761904760,
    /* pos  4 */ 0.3636363640, 0.7200000000, 0.5000000000, 1.0000000000,  0.6363636360, 0.6000000000, 1.0000000000, 0.5714285710,  0.3000000000, 1.0000000000, 0.7857142860, 0.8666666670,  1.0000000000, 0.6400000000, 1.0000000000, 0.5000000000,
    /* pos  5 */ 0.7142857140, 0.7142857140, 0.4545454550, 1.0000000000,  0.9285714290, 0.5000000000, 1.0000000000, 0.9285714290,  0.6666666670, 1.0000000000, 0.6818181820, 1.0000000000,  1.0000000000, 0.5714285710, 0.9090909090, 0.8666666670,
    /* pos  6 */ 0.4375000000, 0.7058823530, 0.4375000000, 1.0000000000,  0.8125000000, 0.4705882350, 1.0000000000, 0.7500000000,  0.5714285710, 1.0000000000, 0.6875000000, 1.0000000000,  1.0000000000, 0.5882352940, 0.6875000000, 0.8750000000,
    /* pos  7 */ 0.4285714290, 0.7333333330, 0.4285714290, 1.0000000000,  0.8750000000, 0.6428571430, 1.0000000000, 0.6500000000,  0.6250000000, 1.0000000000, 0.6153846150, 1.0000000000,  1.0000000000, 0.7333333330, 1.0000000000, 0.8000000000,
    /* pos  8 */ 0.6000000000, 0.6666666670, 0.5714285710, 1.0000000000,  0.8750000000, 0.6190476190, 1.0000000000, 0.8571428570,  0.5333333330, 1.0000000000, 0.5384615380, 0.6428571430,  1.0000000000, 0.6190476190, 0.9230769230, 0.9285714290,
    /* pos  9 */ 0.8823529410, 0.5555555560, 0.3333333330, 1.0000000000,  0.9411764710, 0.3888888890, 1.0000000000, 0.8666666670,  0.8125000000, 1.0000000000, 0.4000000000, 0.9333333330,  1.0000000000, 0.5000000000, 0.5333333330, 0.8571428570,
    /* pos 10 */ 0.3076923080, 0.6500000000, 0.4000000000, 1.0000000000,  0.3076923080, 0.2500000000, 1.0000000000, 0.7500000000,  0.3846153850, 1.0000000000, 0.4285714290, 1.0000000000,  1.0000000000, 0.4000000000, 0.6666666670, 0.7500000000,
    /* pos 11 */ 0.3333333330, 0.7222222220, 0.2631578950, 1.0000000000,  0.5384615380, 0.4444444440, 1.0000000000, 0.7142857140,  0.3846153850, 1.0000000000, 0.5294117650, 0.9333333330,  1.0000000000, 0.5000000000, 0.9473684210, 0.8000000000,
    /* pos 12 */ 0.3000000000, 0.6521739130, 0.2105263160, 1.0000000000,  0.7000000000, 0.1363636360, 1.0000000000, 0.3846153850,  0.3000000000, 1.0000000000, 0.4210526320, 0.9230769230,  1.0000000000, 0.2608695650, 0.7894736840, 0.6923076920,
    /* pos 13 */ 0.5333333330, 0.4666666670, 0.2142857140, 1.0000000000,  0.7333333330, 0.0000000000, 1.0000000000, 0.3500000000,  0.2666666670, 1.0000000000, 0.4285714290, 0.7500000000,  1.0000000000, 0.0000000000, 0.2857142860, 0.6190476190,
    /* pos 14 */ 0.2000000000, 0.6500000000, 0.2727272730, 1.0000000000,  0.0666666670, 0.0500000000, 1.0000000000, 0.2222222220,  0.1428571430, 1.0000000000, 0.2727272730, 0.9411764710,  1.0000000000, 0.0500000000, 0.2727272730, 0.5789473680,
    /* pos 15 */ 0.0000000000, 0.1923076920, 0.0000000000, 1.0000000000,  0.3076923080, 0.1538461540, 1.0000000000, 1.0000000000,  0.0000000000, 1.0000000000, 0.0000000000, 1.0000000000,  1.0000000000, 0.3461538460, 0.6666666670, 0.9090909090,
    /* pos 16 */ 0.1333333330, 0.1764705880, 0.1764705880, 1.0000000000,  0.4666666670, 0.0588235290, 1.0000000000, 0.4666666670,  0.2500000000, 1.0000000000, 0.2352941180, 0.9333333330,  1.0000000000, 0.1176470590, 0.7058823530, 0.5333333330,
    /* pos 17 */ 0.5000000000, 0.4000000000, 0.1904761900, 1.0000000000,  0.6428571430, 0.1333333330, 1.0000000000, 0.5384615380,  0.6666666670, 1.0000000000, 0.4761904760, 0.6923076920,  1.0000000000, 0.3333333330, 0.4285714290, 0.6666666670,
    /* pos 18 */ 0.5384615380, 0.3750000000, 0.2068965520, 1.0000000000,  0.4615384620, 0.1250000000, 1.0000000000, 0.4285714290,  0.6666666670, 1.0000000000, 0.4482758620, 0.7142857140,  1.0000000000, 0.2500000000, 0.2758620690, 0.2857142860,
    /* pos 19 */ 0.6000000000, 0.7647058820, 0.2272727270, 1.0000000000,  0.3000000000, 0.0588235290, 1.0000000000, 0.5000000000,  0.7000000000, 1.0000000000, 0.4285714290, 0.9375000000,  1.0000000000, 0.1764705880, 0.0909090910, 0.5625000000,
};

/* lib/otScorePenalties.cpp:6529-6546, index (nt1<<2)|nt2 of the PAM. */
static const double cfd_pam[16] = {
    /* A_ */ 0.0000000000, 0.0000000000, 0.2592592590, 0.0000000000,
    /* C_ */ 0.0000000000, 0.0000000000, 0.1071428570, 0.0000000000,
    /* G_ */ 0.0694444440, 0.0222222220, 1.0000000000, 0.0161290320,
    /* T_ */ 0.0000000000, 0.0000000000, 0.0389610390, 0.0000000000,
};

void or_cfd_tables(double *pos320, double *pam16)
{
    if (pos320)
        memcpy(pos320, cfd_pos, sizeof(cfd_pos));
    if (pam16)
        memcpy(pam16, cfd_pam, sizeof(cfd_pam));
}

/* --------------------------------------------------------------- scoring */

typedef struct {
    uint64_t *keys;
    double *vals;
    size_t n;
} mit_map;

/* The reference looks MIT up with phmap .at() (ISSLScoringModule.cpp:296);
 * a sorted array with binary search gives the same value for every key. */
static double mit_lookup(const mit_map *m, uint64_t key)
{
    size_t lo = 0, hi = m->n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (m->keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < m->n && m->keys[lo] == key)
        return m->vals[lo];
    abort(); /* .at() would throw: unreachable for 1..4 mismatches */
}

/* ISSLScoringModule.cpp:301-363: product of position penalties where the
 * guide and the site differ, starting from the NGG PAM penalty. */
static double cfd_site(uint64_t q, uint64_t ot, uint64_t dist)
{
    if (dist == 0)
        return 1;
    double s = cfd_pam[10];
    for (unsigned p = 0; p < 20; ++p) {
        unsigned g = (unsigned)((q >> (2 * p)) & 3);
        unsigned o = (unsigned)((ot >> (2 * p)) & 3);
        if (g != o)
            s *= cfd_pos[(p << 4) | (g << 2) | (o ^ 3)];
    }
    return s;
}

/* Relative distance of the exit tests' operands from the cut
 * (ISSLScoringModule.cpp:365-395): how far the running sums are from
 * flipping a test.  0 when an operand equals the cut; +inf for an infinite
 * cut (threshold 0) and for a cut of 0 (threshold 100): a sum of nonnegative
 * terms is 0 or positive whatever order adds them. */
static double exit_distance(int method, double m, double c, double cut)
{
    if (isinf(cut) || cut == 0.0)
        return HUGE_VAL;
    double d;
    switch (method) {
    case OR_MIT_AND_CFD:
    case OR_MIT_OR_CFD: d = fmin(fabs(m - cut), fabs(c - cut)); break;
    case OR_AVG:        d = fabs((m + c) / 2.0 - cut); break;
    case OR_MIT:        d = fabs(m - cut); break;
    default:            d = fabs(c - cut); break;
    }
    return d / fabs(cut);
}

/* Exit tests of ISSLScoringModule.cpp:365-395. */
static int exit_now(int method, double m, double c, double cut)
{
    switch (method) {
    case OR_MIT_AND_CFD: return m > cut && c > cut;
    case OR_MIT_OR_CFD:  return m > cut || c > cut;
    case OR_AVG:         return ((m + c) / 2.0) > cut;
    case OR_MIT:         return m > cut;
    default:             return c > cut;
    }
}

int or_score(const uint64_t *sigs, uint64_t n_sites, uint64_t seq_len,
             const uint64_t *slice_masks, uint64_t n_slices,
             const uint64_t *list_sizes, const uint64_t *entries,
             const uint64_t *guides, uint64_t n_guides,
             int method, double threshold, int threads,
             double *mit_out, double *cfd_out, uint64_t *occ_out,
             uint64_t *site_out, char *pass_out,
             double *mit_sum_out, double *cfd_sum_out, uint64_t *distinct,
             double *margin_out)
{
    if (method < 0 || method > 4)
        return -1;
    const int calc_mit = method != OR_CFD;
    const int calc_cfd = method != OR_MIT;

    /* Slice positions and per-list start offsets
     * (ISSLScoringModule.cpp:100-147,181-198). */
    uint64_t *npos = (uint64_t *)calloc(n_slices ? n_slices : 1, sizeof(uint64_t));
    uint64_t *pos = (uint64_t *)calloc((n_slices ? n_slices : 1) * 64, sizeof(uint64_t));
    uint64_t *size_base = (uint64_t *)calloc(n_slices + 1, sizeof(uint64_t));
    for (uint64_t s = 0; s < n_slices; ++s) {
        for (uint64_t j = 0; j < seq_len && j < 64; ++j)
            if (slice_masks[s] & (1ULL << j))
                pos[s * 64 + npos[s]++] = j;
        if (npos[s] > 16) {
            free(npos); free(pos); free(size_base);
            return -1;
        }
        size_base[s + 1] = size_base[s] + (1ULL << (2 * npos[s]));
    }
    const uint64_t n_lists = size_base[n_slices];
    uint64_t *list_start = (uint64_t *)malloc((n_lists ? n_lists : 1) * sizeof(uint64_t));
    for (uint64_t s = 0; s < n_slices; ++s) {
        uint64_t off = s * n_sites;
        for (uint64_t k = size_base[s]; k < size_base[s + 1]; ++k) {
            list_start[k] = off;
            off += list_sizes[k];
        }
    }

    mit_map mm;
    mm.n = or_mit_table(NULL, NULL);
    mm.keys = (uint64_t *)malloc(mm.n * sizeof(uint64_t));
    mm.vals = (double *)malloc(mm.n * sizeof(double));
    or_mit_table(mm.keys, mm.vals);

    const uint64_t words = n_sites / 64 + 1; /* ISSLScoringModule.cpp:162 */
    if (threads > 0)
        omp_set_num_threads(threads);

#pragma omp parallel
    {
        uint64_t *seen = (uint64_t *)calloc(words, sizeof(uint64_t));
#pragma omp for schedule(static)
        for (int64_t gi = 0; gi < (int64_t)n_guides; ++gi) {
            const uint64_t q = guides[gi];
            double tot_mit = 0.0, tot_cfd = 0.0;
            uint64_t occ_sum = 0, site_sum = 0;
            const double cut = (10000.0 - threshold * 100) / threshold;
            double margin = HUGE_VAL;
            int go_on = 1;
            for (uint64_t s = 0; s < n_slices && go_on; ++s) {
                uint64_t key = 0;
                for (uint64_t j = 0; j < npos[s]; ++j)
                    key |= ((q >> (pos[s * 64 + j] * 2)) & 3ULL) << (j * 2);
                const uint64_t li = size_base[s] + key;
                const uint64_t len = list_sizes[li];
                const uint64_t *list = entries + list_start[li];
                for (uint64_t e = 0; e < len; ++e) {
                    const uint64_t id = list[e] & 0xFFFFFFFFULL;
                    const uint32_t occ = (uint32_t)(list[e] >> 32);
                    uint64_t *w = &seen[id / 64];
                    const uint64_t bit = 1ULL << (id % 64);
                    if (*w & bit)
                        continue;
                    *w |= bit;
                    const uint64_t x = q ^ sigs[id];
                    const uint64_t fold = ((x >> 1) & FOLD_LO) | (x & FOLD_LO);
                    const uint64_t dist = (uint64_t)__builtin_popcountll(fold);
                    if (dist > 4)
                        continue;
                    occ_sum += occ;
                    site_sum += 1;
                    if (distinct) {
#pragma omp atomic
                        distinct[id / 64] |= bit;
                    }
                    if (calc_mit && dist > 0)
                        tot_mit += mit_lookup(&mm, fold) * (double)occ;
                    if (calc_cfd)
                        tot_cfd += cfd_site(q, sigs[id], dist) * (double)occ;
                    if (margin_out) {
                        const double d = exit_distance(method, tot_mit, tot_cfd, cut);
                        if (d < margin)
                            margin = d;
                    }
                    if (exit_now(method, tot_mit, tot_cfd, cut)) {
                        go_on = 0;
                        break;
                    }
                }
            }
            const double smit = 10000.0 / (100.0 + tot_mit);
            const double scfd = 10000.0 / (100.0 + tot_cfd);
            int reject;
            switch (method) {
            case OR_MIT_AND_CFD: reject = smit < threshold && scfd < threshold; break;
            case OR_MIT_OR_CFD:  reject = smit < threshold || scfd < threshold; break;
            case OR_AVG:         reject = (smit + scfd) / 2 < threshold; break;
            case OR_MIT:         reject = smit < threshold; break;
            default:             reject = scfd < threshold; break;
            }
            if (mit_out) mit_out[gi] = calc_mit ? smit : -1;
            if (cfd_out) cfd_out[gi] = calc_cfd ? scfd : -1;
            if (occ_out) occ_out[gi] = occ_sum;
            if (site_out) site_out[gi] = site_sum;
            if (pass_out) pass_out[gi] = reject ? '0' : '1';
            if (mit_sum_out) mit_sum_out[gi] = tot_mit;
            if (cfd_sum_out) cfd_sum_out[gi] = tot_cfd;
            if (margin_out) margin_out[gi] = margin;
            memset(seen, 0, words * sizeof(uint64_t)); /* ISSLScoringModule.cpp:461 */
        }
        free(seen);
    }
    free(mm.keys); free(mm.vals); free(list_start);
    free(npos); free(pos); free(size_base);
    return 0;
}

/* ------------------------------------------------------------ brute force */

uint64_t or_candidates_brute(const uint64_t *sigs, uint64_t n_sites,
                             uint64_t seq_len, const uint64_t *slice_masks,
                             uint64_t n_slices, uint64_t q,
                             uint32_t *ids_out, uint64_t cap)
{
    uint64_t fm[64];
    uint64_t ns = n_slices < 64 ? n_slices : 64;
    for (uint64_t s = 0; s < ns; ++s) {
        fm[s] = 0;
        for (uint64_t j = 0; j < seq_len && j < 32; ++j)
            if (slice_masks[s] & (1ULL << j))
                fm[s] |= 1ULL << (2 * j);
    }
    uint64_t n = 0;
    for (uint64_t id = 0; id < n_sites; ++id) {
        const uint64_t x = q ^ sigs[id];
        const uint64_t fold = ((x >> 1) & FOLD_LO) | (x & FOLD_LO);
        if (__builtin_popcountll(fold) > 4)
            continue;
        int listed = 0;
        for (uint64_t s = 0; s < ns && !listed; ++s)
            listed = (fold & fm[s]) == 0;
        if (!listed)
            continue;
        if (n < cap)
            ids_out[n] = (uint32_t)id;
        ++n;
    }
    return n;
}

/* ---------------------------------------------------------------- extract */

/* Base of a scanned sequence: the reference upper-cases every line
 * (ExtractOfftargets.cpp:88-98) and its patterns admit only A, C, G, T. */
static int or_base(char c)
{
    if (c >= 'a' && c <= 'z')
        c = (char)(c - 'a' + 'A');
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
    }
}

uint64_t or_extract(const char *seq, uint64_t len, char *out, uint64_t cap)
{
    static const char bases[4] = {'A', 'C', 'G', 'T'};
    uint64_t n = 0;
    /* forward: (?=([ACG][ACGT]{19}[ACGT][AG]G)), first 20 characters
     * (ExtractOfftargets.cpp:16,127-141) */
    for (uint64_t i = 0; i + 23 <= len; ++i) {
        int ok = 1;
        for (int j = 0; j < 23 && ok; ++j)
            ok = or_base(seq[i + j]) >= 0;
        if (!ok)
            continue;
        if (or_base(seq[i]) == 3)
            continue;
        const int b21 = or_base(seq[i + 21]);
        if (!(b21 == 0 || b21 == 2) || or_base(seq[i + 22]) != 2)
            continue;
        if (n < cap)
            for (int j = 0; j < 20; ++j)
                out[n * 20 + j] = bases[or_base(seq[i + j])];
        ++n;
    }
    /* reverse: (?=(C[CT][ACGT][ACGT]{19}[TGC])), reverse complement of the
     * 23 characters (lib/util.cpp:50-63), first 20
     * (ExtractOfftargets.cpp:17,142-156) */
    for (uint64_t i = 0; i + 23 <= len; ++i) {
        int ok = 1;
        for (int j = 0; j < 23 && ok; ++j)
            ok = or_base(seq[i + j]) >= 0;
        if (!ok)
            continue;
        const int b1 = or_base(seq[i + 1]);
        if (or_base(seq[i]) != 1 || !(b1 == 1 || b1 == 3) || or_base(seq[i + 22]) == 0)
            continue;
        if (n < cap)
            for (int j = 0; j < 20; ++j)
                out[n * 20 + j] = bases[3 - or_base(seq[i + 22 - j])];
        ++n;
    }
    return n;
}
