// common.hpp -- status codes, the thread-local error message and small
// encoding helpers shared by the host (g++) and HIP (hipcc) units of
// libissl_gpu.so.
#pragma once

#include <cstdint>
#include <string>

#include "issl_gpu.h"

namespace issl {

// Records msg as this thread's last error and returns code.
int fail(int code, const std::string &msg);
// Clears the thread's last error; returns ISSL_OK.
int ok();

// Fold of q ^ site: bit 2j set iff position j differs
// (lib/ISSLScoringModule.cpp:286-290).
constexpr uint64_t kFoldLo = 0x5555555555555555ULL;
inline uint64_t fold_of(uint64_t x) { return ((x >> 1) & kFoldLo) | (x & kFoldLo); }

// Number of MIT table entries: fold masks with 1..4 of 20 positions.
constexpr int kMitEntries = 6195;
// First combinadic rank of each mismatch count k (index k = 1..4).
constexpr int kMitBase[5] = {0, 0, 20, 210, 1350};
// MIT factor table (mit_factor_table): kMitOneMinus values 1 - weight[p] by
// position (1.0 from position 20 on), then for k = 0..4 and span = 0..19 the
// pair (t2, t3) of the closed form at [kMitOneMinus + 2 (20 k + span)].
constexpr int kMitOneMinus = 33;
constexpr int kMitFactorEntries = kMitOneMinus + 2 * 5 * 20;

// Slice widths are limited to 16 positions: the reference builds keys in a
// uint32 (ISSLCreateIndex.cpp:458).
constexpr int kMaxWidth = 16;
// Slice count supported by the device kernels.
constexpr int kMaxSlices = 64;

// Positions of a slice mask below seq_len, ascending
// (ISSLScoringModule.cpp:101-115).
int mask_positions(uint64_t mask, uint64_t seq_len, uint8_t *pos_out);

// Number of posting entries whose id is >= n_sites (host, OpenMP).  The
// device kernels gather sig[id], so an index is only uploaded when this is 0.
uint64_t count_bad_ids(const uint64_t *entries, uint64_t count, uint64_t n_sites);
// Signatures with bits at or above bit 40 (more than 20 positions).
uint64_t count_wide_sigs(const uint64_t *sigs, uint64_t n);

const char *last_error_cstr();

// MIT values in combinadic order (rank of the mismatch position set) and the
// matching fold masks; filled from the closed form on first use.
const double *mit_by_rank();
const uint64_t *mit_keys_by_rank();
// The MIT closed form as factors (kMitFactorEntries values), and whether
// they reproduce every value of mit_by_rank() bit for bit.
const double *mit_factor_table();
bool mit_factors_exact();
// CFD tables: 320 position penalties, 16 PAM penalties.
const double *cfd_pos_table();
const double *cfd_pam_table();

}  // namespace issl
