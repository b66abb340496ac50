// index_host.cpp -- .issl construction and file I/O on the host (OpenMP).
//
// Builder semantics follow ISSLCreateIndex/ISSLCreateIndex.cpp:396-484:
// a site's id is its position in the distinct-site array, each slice groups
// ids by slice key, ids ascend within a key, entries pack occ<<32 | id.
// The grouping here is a stable parallel counting sort instead of the
// reference's vector-of-vectors.
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "issl_index.h"

namespace issl {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int ok() {
    g_last_error.clear();
    return ISSL_OK;
}

const char *last_error_cstr() { return g_last_error.c_str(); }

}  // namespace issl

using namespace issl;

extern "C" const char *issl_last_error(void) { return last_error_cstr(); }

uint64_t issl::count_bad_ids(const uint64_t *entries, uint64_t count, uint64_t n_sites) {
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(count); ++i)
        bad += (entries[i] & 0xFFFFFFFFULL) >= n_sites ? 1 : 0;
    return bad;
}

uint64_t issl::count_wide_sigs(const uint64_t *sigs, uint64_t n) {
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) bad += (sigs[i] >> 40) != 0 ? 1 : 0;
    return bad;
}

extern "C" uint64_t issl_list_count(const uint64_t *slice_masks, uint64_t n_slices,
                                    uint64_t seq_len) {
    uint64_t total = 0;
    for (uint64_t s = 0; s < n_slices; ++s) {
        uint8_t pos[64];
        const int w = mask_positions(slice_masks[s], seq_len, pos);
        if (w > kMaxWidth) return 0;
        total += 1ULL << (2 * w);
    }
    return total;
}

extern "C" int issl_build_lists(const uint64_t *sigs, const uint32_t *occ,
                                uint64_t n_sites, uint64_t seq_len,
                                const uint64_t *slice_masks, uint64_t n_slices,
                                uint64_t *list_sizes, uint64_t *entries,
                                int threads) {
    if ((n_sites && (!sigs || !occ || !entries)) || (n_slices && !slice_masks) || !list_sizes)
        return fail(ISSL_ERR_ARG, "issl_build_lists: null buffer");
    if (n_sites >= (1ULL << 32))
        return fail(ISSL_ERR_ARG, "issl_build_lists: ids are 32-bit (ISSLCreateIndex.cpp:466)");
    if (n_slices && issl_list_count(slice_masks, n_slices, seq_len) == 0)
        return fail(ISSL_ERR_ARG, "issl_build_lists: slice wider than 16 positions");
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    uint64_t size_base = 0;
    for (uint64_t s = 0; s < n_slices; ++s) {
        uint8_t pos[64];
        const int w = mask_positions(slice_masks[s], seq_len, pos);
        const uint64_t nkeys = 1ULL << (2 * w);
        // Per-thread histograms over contiguous id ranges make the scatter
        // stable: thread t writes its ids after every lower thread's ids of
        // the same key.  Fewer threads when the histograms would be large.
        int t_use = nt;
        while (t_use > 1 && static_cast<double>(t_use) * nkeys * 8 > 2.0e9) --t_use;
        std::vector<uint64_t> hist(static_cast<size_t>(t_use) * nkeys, 0);
        std::vector<uint32_t> keys(n_sites);
        uint64_t *out = entries + s * n_sites;
#pragma omp parallel num_threads(t_use)
        {
            const int t = omp_get_thread_num();
            const uint64_t lo = n_sites * t / t_use, hi = n_sites * (t + 1) / t_use;
            uint64_t *h = hist.data() + static_cast<size_t>(t) * nkeys;
            for (uint64_t id = lo; id < hi; ++id) {
                uint32_t key = 0;
                for (int j = 0; j < w; ++j)
                    key |= static_cast<uint32_t>((sigs[id] >> (2 * pos[j])) & 3ULL) << (2 * j);
                keys[id] = key;
                ++h[key];
            }
#pragma omp barrier
#pragma omp single
            {
                uint64_t run = 0;
                for (uint64_t k = 0; k < nkeys; ++k) {
                    uint64_t sz = 0;
                    for (int tt = 0; tt < t_use; ++tt) {
                        uint64_t &c = hist[static_cast<size_t>(tt) * nkeys + k];
                        const uint64_t cnt = c;
                        c = run;
                        run += cnt;
                        sz += cnt;
                    }
                    list_sizes[size_base + k] = sz;
                }
            }
            for (uint64_t id = lo; id < hi; ++id)
                out[h[keys[id]]++] = (static_cast<uint64_t>(occ[id]) << 32) | id;
        }
        size_base += nkeys;
    }
    return ok();
}

namespace {

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

bool write_words(FILE *f, const uint64_t *p, uint64_t n) {
    const uint64_t step = 1ULL << 24;
    for (uint64_t i = 0; i < n; i += step) {
        const uint64_t c = n - i < step ? n - i : step;
        if (fwrite(p + i, sizeof(uint64_t), c, f) != c) return false;
    }
    return true;
}

bool read_words(FILE *f, uint64_t *p, uint64_t n) {
    const uint64_t step = 1ULL << 24;
    for (uint64_t i = 0; i < n; i += step) {
        const uint64_t c = n - i < step ? n - i : step;
        if (fread(p + i, sizeof(uint64_t), c, f) != c) return false;
    }
    return true;
}

}  // namespace

extern "C" int issl_write_index(const char *path, uint64_t n_sites, uint64_t seq_len,
                                uint64_t n_slices, const uint64_t *sigs,
                                const uint64_t *slice_masks, const uint64_t *list_sizes,
                                const uint64_t *entries) {
    if (!path) return fail(ISSL_ERR_ARG, "issl_write_index: null path");
    const uint64_t n_lists = issl_list_count(slice_masks, n_slices, seq_len);
    if (n_slices && n_lists == 0) return fail(ISSL_ERR_ARG, "issl_write_index: slice too wide");
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) return fail(ISSL_ERR_IO, std::string("cannot open for writing: ") + path);
    const uint64_t hdr[3] = {n_sites, seq_len, n_slices};
    bool good = write_words(fh.f, hdr, 3) && write_words(fh.f, sigs, n_sites) &&
                write_words(fh.f, slice_masks, n_slices);
    uint64_t base = 0;
    for (uint64_t s = 0; good && s < n_slices; ++s) {
        uint8_t pos[64];
        const uint64_t nk = 1ULL << (2 * mask_positions(slice_masks[s], seq_len, pos));
        good = write_words(fh.f, list_sizes + base, nk) &&
               write_words(fh.f, entries + s * n_sites, n_sites);
        base += nk;
    }
    if (!good) return fail(ISSL_ERR_IO, std::string("write failed: ") + path);
    return ok();
}

extern "C" int issl_read_header(const char *path, uint64_t *n_sites, uint64_t *seq_len,
                                uint64_t *n_slices, uint64_t *n_lists) {
    if (!path) return fail(ISSL_ERR_ARG, "issl_read_header: null path");
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return fail(ISSL_ERR_IO, std::string("Error reading index: could not open file ") + path);
    uint64_t hdr[3];
    if (!read_words(fh.f, hdr, 3)) return fail(ISSL_ERR_FORMAT, "Error reading index: header invalid");
    if (hdr[0] >= (1ULL << 32) || hdr[2] > 4096 || hdr[1] > 64)
        return fail(ISSL_ERR_FORMAT, "Error reading index: header out of range");
    std::vector<uint64_t> masks(hdr[2]);
    if (fseek(fh.f, static_cast<long>(24 + 8 * hdr[0]), SEEK_SET) != 0 ||
        !read_words(fh.f, masks.data(), hdr[2]))
        return fail(ISSL_ERR_FORMAT, "Error reading index: slice masks missing");
    const uint64_t nl = issl_list_count(masks.data(), hdr[2], hdr[1]);
    if (hdr[2] && nl == 0) return fail(ISSL_ERR_FORMAT, "Error reading index: slice wider than 16");
    // The file must hold exactly header + sigs + masks + sizes + S*N entries.
    if (fseek(fh.f, 0, SEEK_END) != 0) return fail(ISSL_ERR_IO, "seek failed");
    const long long fsz = ftell(fh.f);
    const unsigned long long want = 8ULL * (3 + hdr[0] + hdr[2] + nl + hdr[2] * hdr[0]);
    if (fsz < 0 || static_cast<unsigned long long>(fsz) != want)
        return fail(ISSL_ERR_FORMAT, "Error reading index: file size " + std::to_string(fsz) +
                                         " does not match header (" + std::to_string(want) + ")");
    if (n_sites) *n_sites = hdr[0];
    if (seq_len) *seq_len = hdr[1];
    if (n_slices) *n_slices = hdr[2];
    if (n_lists) *n_lists = nl;
    return ok();
}

extern "C" int issl_read_index(const char *path, uint64_t *sigs, uint64_t *slice_masks,
                               uint64_t *list_sizes, uint64_t *entries) {
    uint64_t n, L, S, nl;
    const int rc = issl_read_header(path, &n, &L, &S, &nl);
    if (rc != ISSL_OK) return rc;
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return fail(ISSL_ERR_IO, std::string("cannot open ") + path);
    uint64_t hdr[3];
    bool good = read_words(fh.f, hdr, 3) && read_words(fh.f, sigs, n) &&
                read_words(fh.f, slice_masks, S);
    uint64_t base = 0;
    for (uint64_t s = 0; good && s < S; ++s) {
        uint8_t pos[64];
        const uint64_t nk = 1ULL << (2 * mask_positions(slice_masks[s], L, pos));
        good = read_words(fh.f, list_sizes + base, nk) && read_words(fh.f, entries + s * n, n);
        if (good) {
            uint64_t tot = 0;
            for (uint64_t k = 0; k < nk; ++k) tot += list_sizes[base + k];
            if (tot != n)
                return fail(ISSL_ERR_FORMAT, "Error reading index: slice " + std::to_string(s) +
                                                 " list sizes do not sum to N");
        }
        base += nk;
    }
    if (!good) return fail(ISSL_ERR_FORMAT, "Error reading index: truncated");
    return ok();
}

extern "C" int issl_parse_masks(const char *text, uint64_t len, uint64_t seq_len,
                                uint64_t *masks, uint64_t cap, uint64_t *n_out) {
    if (!text || !n_out) return fail(ISSL_ERR_ARG, "issl_parse_masks: null buffer");
    uint64_t n = 0, i = 0;
    while (i < len) {
        uint64_t e = i;
        while (e < len && text[e] != '\n') ++e;
        uint64_t lend = e;
        if (lend > i && text[lend - 1] == '\r') --lend;
        if (lend > i) {
            uint64_t m = 0;
            for (uint64_t j = 0; j < seq_len && i + j < lend; ++j)
                if (text[i + j] == '1') m |= 1ULL << j;
            if (n < cap && masks) masks[n] = m;
            ++n;
        }
        i = e + 1;
    }
    *n_out = n;
    return ok();
}

namespace {
unsigned code_of(char c) {
    switch (c) {
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return 0;
    }
}
}  // namespace

extern "C" int issl_count_sorted_sites(const char *text, uint64_t len, uint64_t seq_len,
                                       uint64_t line_len, uint64_t *sigs, uint32_t *occ,
                                       uint64_t cap, uint64_t *n_out) {
    if (!text || !n_out || line_len < seq_len || seq_len == 0 || seq_len > 32)
        return fail(ISSL_ERR_ARG, "issl_count_sorted_sites: bad arguments");
    if (len % line_len != 0)
        return fail(ISSL_ERR_FORMAT,
                    "offtargetSites file is not a multiple of the expected line length (" +
                        std::to_string(line_len) + ")");
    const uint64_t lines = len / line_len;
    uint64_t n = 0, i = 0;
    while (i < lines) {
        const char *p = text + i * line_len;
        uint32_t c = 1;
        while (i + c < lines && std::memcmp(p, p + c * line_len, seq_len) == 0) ++c;
        if (n < cap) {
            uint64_t sig = 0;
            for (uint64_t j = 0; j < seq_len; ++j) sig |= static_cast<uint64_t>(code_of(p[j])) << (2 * j);
            if (sigs) sigs[n] = sig;
            if (occ) occ[n] = c;
        }
        ++n;
        i += c;
    }
    *n_out = n;
    return ok();
}
