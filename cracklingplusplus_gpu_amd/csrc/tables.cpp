// tables.cpp -- score tables of the MI355X scorer.
//
// MIT: generated from the closed form of Hsu et al. 2013 as the reference's
// index builder states it (ISSLCreateIndex/ISSLCreateIndex.cpp:235-282), in
// the same operation order, so every value equals the reference's
// precalculatedMITScores entry (lib/otScorePenalties.cpp:3-6199) bit for bit.
// Entries are stored by the combinadic rank of the mismatch position set,
// the order the device kernels index.
//
// CFD: Doench et al. 2016 penalties, the constants of
// lib/otScorePenalties.cpp:6201-6546 (published data; kept identical because
// the scores depend on every digit).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "issl_index.h"

namespace issl {
namespace {

// Position weights of the MIT formula (ISSLCreateIndex.cpp:239).
const double kMitWeight[20] = {0.0,   0.0,   0.014, 0.0,   0.0,   0.395, 0.317,
                               0.0,   0.389, 0.079, 0.445, 0.508, 0.613, 0.851,
                               0.732, 0.828, 0.615, 0.804, 0.685, 0.583};

double mit_closed_form(const int *p, int k) {
    double t1 = 1.0;
    for (int i = 0; i < k; ++i) t1 = t1 * (1.0 - kMitWeight[p[i]]);
    double gap = 0.0;
    if (k == 1) {
        gap = 19.0;
    } else {
        for (int i = 0; i + 1 < k; ++i) gap += static_cast<double>(p[i + 1] - p[i]);
        gap = gap / static_cast<double>(k - 1);
    }
    const double t2 = 1.0 / ((19.0 - gap) / 19.0 * 4.0 + 1);
    const double t3 = 1.0 / static_cast<double>(static_cast<uint64_t>(k) * k);
    return t1 * t2 * t3 * 100;
}

uint64_t binom(int n, int r) {
    if (r < 0 || n < r) return 0;
    uint64_t v = 1;
    for (int i = 1; i <= r; ++i) v = v * static_cast<uint64_t>(n - r + i) / i;
    return v;
}

struct MitTable {
    double val[kMitEntries];
    uint64_t key[kMitEntries];
    MitTable() {
        // rank(p1<..<pk) = base[k] + sum_i C(p_i, i): the combinatorial number
        // system, dense over each k.
        for (uint32_t bits = 1; bits < (1u << 20); ++bits) {
            const int k = __builtin_popcount(bits);
            if (k > 4) continue;
            int p[4];
            int m = 0;
            uint64_t fold = 0;
            for (int j = 0; j < 20; ++j)
                if ((bits >> j) & 1u) {
                    p[m++] = j;
                    fold |= 1ULL << (2 * j);
                }
            uint64_t rank = kMitBase[k];
            for (int i = 0; i < k; ++i) rank += binom(p[i], i + 1);
            val[rank] = mit_closed_form(p, k);
            key[rank] = fold;
        }
    }
};

const MitTable &mit_table() {
    static const MitTable t;
    return t;
}

// The closed form as factors the kernels multiply: one minus the position
// weight (1.0 past position 19, so absent mismatches multiply exactly), and
// per mismatch count k and span (last minus first position) the second and
// third terms.  mit_closed_form's t1 starts at 1.0 and multiplies the
// factors in position order, so om[p0]*om[p1]*om[p2]*om[p3] is the same
// rounded value; t2 depends on the positions only through the span (the sum
// of the gaps is the span, exactly), t3 only on k.  mit_factors_exact()
// checks that ((t1 * t2) * t3) * 100 from these factors equals every table
// value bit for bit.
struct MitFactors {
    double f[kMitFactorEntries];
    bool exact = true;
    MitFactors() {
        for (int p = 0; p < kMitOneMinus; ++p) f[p] = p < 20 ? 1.0 - kMitWeight[p] : 1.0;
        double *t23 = f + kMitOneMinus;
        for (int k = 0; k <= 4; ++k)
            for (int span = 0; span < 20; ++span) {
                double t2 = 0.0, t3 = 0.0;
                if (k >= 1) {
                    const double gap = k == 1 ? 19.0 : static_cast<double>(span) / static_cast<double>(k - 1);
                    t2 = 1.0 / ((19.0 - gap) / 19.0 * 4.0 + 1);
                    t3 = 1.0 / static_cast<double>(static_cast<uint64_t>(k) * k);
                }
                t23[2 * (k * 20 + span)] = t2;
                t23[2 * (k * 20 + span) + 1] = t3;
            }
        const MitTable &t = mit_table();
        for (int r = 0; r < kMitEntries; ++r) {
            int pos[4] = {32, 32, 32, 32}, k = 0;
            for (int j = 0; j < 20; ++j)
                if ((t.key[r] >> (2 * j)) & 1ULL) pos[k++] = j;
            double t1 = f[pos[0]] * f[pos[1]];
            t1 = t1 * f[pos[2]];
            t1 = t1 * f[pos[3]];
            const int span = pos[k - 1] - pos[0];
            const double v = ((t1 * t23[2 * (k * 20 + span)]) * t23[2 * (k * 20 + span) + 1]) * 100.0;
            if (std::memcmp(&v, &t.val[r], sizeof v) != 0) exact = false;
        }
    }
};

const MitFactors &mit_factors() {
    static const MitFactors t;
    return t;
}

// One row per position; four groups of four: guide nucleotide, then the
// reverse complement of the site nucleotide (index (pos<<4)|(g<<2)|rc(o)).
const double kCfdPos[320] = {
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

// PAM penalties indexed (first<<2)|second nucleotide; NGG is index 10.
const double kCfdPam[16] = {
    /* A_ */ 0.0000000000, 0.0000000000, 0.2592592590, 0.0000000000,
    /* C_ */ 0.0000000000, 0.0000000000, 0.1071428570, 0.0000000000,
    /* G_ */ 0.0694444440, 0.0222222220, 1.0000000000, 0.0161290320,
    /* T_ */ 0.0000000000, 0.0000000000, 0.0389610390, 0.0000000000,
};

}  // namespace

const double *mit_by_rank() { return mit_table().val; }
const double *mit_factor_table() { return mit_factors().f; }
bool mit_factors_exact() { return mit_factors().exact; }
const uint64_t *mit_keys_by_rank() { return mit_table().key; }
const double *cfd_pos_table() { return kCfdPos; }
const double *cfd_pam_table() { return kCfdPam; }

int mask_positions(uint64_t mask, uint64_t seq_len, uint8_t *pos_out) {
    int w = 0;
    for (uint64_t j = 0; j < seq_len && j < 64; ++j)
        if (mask & (1ULL << j)) {
            if (w < 64) pos_out[w] = static_cast<uint8_t>(j);
            ++w;
        }
    return w;
}

}  // namespace issl

extern "C" int issl_score_tables(uint64_t *mit_keys, double *mit_vals,
                                 double *cfd_pos, double *cfd_pam) {
    using namespace issl;
    if (mit_keys || mit_vals) {
        std::vector<std::pair<uint64_t, double>> kv(kMitEntries);
        for (int r = 0; r < kMitEntries; ++r) kv[r] = {mit_keys_by_rank()[r], mit_by_rank()[r]};
        std::sort(kv.begin(), kv.end());
        for (int i = 0; i < kMitEntries; ++i) {
            if (mit_keys) mit_keys[i] = kv[i].first;
            if (mit_vals) mit_vals[i] = kv[i].second;
        }
    }
    if (cfd_pos) std::memcpy(cfd_pos, cfd_pos_table(), 320 * sizeof(double));
    if (cfd_pam) std::memcpy(cfd_pam, cfd_pam_table(), 16 * sizeof(double));
    return ok();
}
