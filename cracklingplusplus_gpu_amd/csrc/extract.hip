// extract.hip -- off-target site extraction on the device (issl_extract.h),
// the step before the index: the reference's ExtractOfftargets
// (ExtractOfftargets/ExtractOfftargets.cpp:19-356).
//
// The reference splits its inputs into one temporary file per sequence,
// runs two lookahead regexes over each, writes the 20-mers to batch files,
// sorts each batch and merges them on disk.  Here:
//   * sequences are packed into pinned chunks (a separator byte between
//     them, so no window spans two) or, when longer than a chunk, cut into
//     pieces overlapping by 22 bases, and copied to HBM;
//   * k_pam_scan tests every start position of a chunk (one thread each,
//     the 23-base windows of a 256-position tile staged in LDS) for the
//     forward pattern [ACG]N{19}N[AG]G and the reverse pattern
//     C[CT]NN{19}[TGC], and stages each site's text-order key (base 0 in
//     the top bits, so numeric order is the reference's string order) in
//     LDS; a staged batch takes its place in HBM with one atomic;
//   * finish sorts the keys with a hipCUB radix sort over their 40 bits and
//     run-length encodes them: the distinct sites in text order and their
//     occurrence counts, which is what ISSLCreateIndex.cpp:396-421 derives
//     from the sorted file;
//   * the sorted file is formatted on the device (21 bytes per site) and
//     written in pieces; or the sites go straight into issl_index_build's
//     device path.
//
// Bound: HBM and the host link.  The scan reads 1 byte per base and writes
// 8 bytes per site; the sort moves each key about ten times.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "device.hpp"
#include "issl_extract.h"

using namespace issl;

struct issl_extractor {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t chunk = 0;            // start positions per launch
    // Two staging buffers used in turn: while the device copies and scans
    // one, the host fills the other.
    HostBuf pin[2];                // packed sequences or one piece: chunk + 23 bytes
    DevBuf d_seq[2];               // device copies
    int cur = 0;                   // buffer being filled
    uint64_t pending = 0;          // bytes in pin[cur]
    bool in_flight = false;        // a scan is queued whose count is not read yet
    HostBuf h_count;               // the counter, copied back after each scan
    uint64_t *d_keys = nullptr;    // text-order keys of every site found
    uint64_t cap = 0, n_keys = 0;
    DevBuf d_count;                // device site counter
    uint64_t bases = 0;
    bool finished = false;
    DevBuf d_sorted, d_uniq, d_occ, d_nruns, d_tmp, d_out;
    uint64_t n_distinct = 0;
};

namespace {

constexpr uint64_t kDefaultChunk = uint64_t(256) << 20;
constexpr int kWin = 23;           // 20 bases + PAM
constexpr int kSiteLen = 20;
constexpr int kScanBlock = 256;
constexpr int kKeyBits = 2 * kSiteLen;
constexpr uint64_t kMaxSites = 0x7FFFFFFFULL;     // hipCUB run-length encode takes int sizes
constexpr uint64_t kFormatSites = uint64_t(32) << 20;

// A, C, G, T in either case -> 0..3; any other byte -> 4 (ends windows).
__device__ __forceinline__ uint32_t base_code(uint8_t c) {
    if (c >= 'a' && c <= 'z') c = static_cast<uint8_t>(c - 'a' + 'A');
    return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : c == 'T' ? 3u : 4u;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t b) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
}

// Keys staged per workgroup before they go to HBM: one atomic reserves the
// space for a whole batch (a per-wave atomic on one counter serializes in
// L2), and the batch is written out coalesced.
constexpr int kStageKeys = 4096;

// Sites starting at positions [0, npos) of seq[0, nbytes) (windows past
// nbytes do not match).  keys must hold *count + 2 * npos entries.
__global__ __launch_bounds__(kScanBlock) void k_pam_scan(const uint8_t *__restrict__ seq,
                                                         uint64_t nbytes, uint64_t npos,
                                                         uint64_t *__restrict__ keys,
                                                         unsigned long long *__restrict__ count) {
    constexpr int kWaves = kScanBlock / 64;
    __shared__ uint8_t code[kScanBlock + kWin - 1];
    __shared__ uint64_t stage[kStageKeys];
    __shared__ uint32_t s_wcnt[kWaves];
    __shared__ unsigned long long s_gbase;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t fill = 0;  // keys staged (uniform)
    auto flush = [&]() {
        if (tid == 0) s_gbase = atomicAdd(count, static_cast<unsigned long long>(fill));
        __syncthreads();
        const unsigned long long g = s_gbase;
        for (uint32_t i = tid; i < fill; i += kScanBlock) keys[g + i] = stage[i];
        __syncthreads();
        fill = 0;
    };
    for (uint64_t base = blockIdx.x * static_cast<uint64_t>(kScanBlock); base < npos;
         base += static_cast<uint64_t>(gridDim.x) * kScanBlock) {
        for (int t = tid; t < kScanBlock + kWin - 1; t += kScanBlock) {
            const uint64_t i = base + t;
            code[t] = static_cast<uint8_t>(i < nbytes ? base_code(seq[i]) : 4u);
        }
        __syncthreads();
        const uint64_t pos = base + tid;
        bool f = false, r = false;
        uint64_t kf = 0, kr = 0;
        if (pos < npos) {
            bool all = true;
            for (int j = 0; j < kWin; ++j) all &= code[tid + j] < 4;
            if (all) {
                const uint32_t c0 = code[tid], c1 = code[tid + 1];
                const uint32_t c21 = code[tid + 21], c22 = code[tid + 22];
                // forward (?=([ACG][ACGT]{19}[ACGT][AG]G)) (ExtractOfftargets.cpp:16)
                f = c0 != 3 && (c21 | 2u) == 2u && c22 == 2;
                // reverse (?=(C[CT][ACGT][ACGT]{19}[TGC])) (ExtractOfftargets.cpp:17)
                r = c0 == 1 && (c1 & 1u) == 1u && c22 != 0;
                if (f || r) {
                    for (int j = 0; j < kSiteLen; ++j) {
                        kf = (kf << 2) | code[tid + j];
                        // reverse complement (lib/util.cpp:50-63): base j of the
                        // site is the complement of base 22 - j of the window
                        kr = (kr << 2) | (3u - code[tid + 22 - j]);
                    }
                }
            }
        }
        const uint64_t bf = __ballot(f), br = __ballot(r);
        const uint32_t nf = static_cast<uint32_t>(__popcll(bf));
        if (lane == 0) s_wcnt[wave] = nf + static_cast<uint32_t>(__popcll(br));
        __syncthreads();
        uint32_t woff = fill, total = 0;
        for (int w = 0; w < kWaves; ++w) {
            woff += w < wave ? s_wcnt[w] : 0u;
            total += s_wcnt[w];
        }
        if (f) stage[woff + lanes_below(bf)] = kf;
        if (r) stage[woff + nf + lanes_below(br)] = kr;
        fill += total;
        __syncthreads();  // code[] and s_wcnt are rewritten next
        if (fill > static_cast<uint32_t>(kStageKeys - 2 * kScanBlock)) flush();
    }
    if (fill) flush();
}

// Text-order key -> signature (position j in bits 2j).
__device__ __forceinline__ uint64_t key_to_sig(uint64_t key) {
    uint64_t s = 0;
    for (int j = 0; j < kSiteLen; ++j) s |= ((key >> (2 * (kSiteLen - 1 - j))) & 3ULL) << (2 * j);
    return s;
}

__global__ void k_key_sig(const uint64_t *__restrict__ keys, uint64_t n, uint64_t *__restrict__ sigs) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        sigs[i] = key_to_sig(keys[i]);
}

// One output line per site: 20 bases and '\n'.
__global__ void k_format(const uint64_t *__restrict__ keys, uint64_t n, char *__restrict__ text) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t k = keys[i];
        char *o = text + i * (kSiteLen + 1);
        for (int j = 0; j < kSiteLen; ++j) o[j] = "ACGT"[(k >> (2 * (kSiteLen - 1 - j))) & 3];
        o[kSiteLen] = '\n';
    }
}

uint32_t grid_for(uint64_t n, uint64_t per_block, uint32_t cap) {
    const uint64_t g = (n + per_block - 1) / per_block;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

int ensure_cap(issl_extractor *ex, uint64_t need) {
    if (need <= ex->cap) return ISSL_OK;
    const uint64_t nc = std::max<uint64_t>(need, 2 * ex->cap);
    void *p = nullptr;
    if (hipMalloc(&p, nc * 8) != hipSuccess)
        return fail(ISSL_ERR_NOMEM, "device allocation of " + std::to_string(nc * 8) +
                                        " bytes for extracted sites failed");
    if (ex->n_keys) {
        ISSL_HIP(hipMemcpyAsync(p, ex->d_keys, ex->n_keys * 8, hipMemcpyDeviceToDevice, ex->stream));
        ISSL_HIP(hipStreamSynchronize(ex->stream));
    }
    if (ex->d_keys) (void)hipFree(ex->d_keys);
    ex->d_keys = static_cast<uint64_t *>(p);
    ex->cap = nc;
    return ISSL_OK;
}

// Copy with several host threads (staging is the scan's critical path).
void par_copy(char *dst, const char *src, uint64_t n) {
    constexpr uint64_t kMinPart = uint64_t(8) << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const uint64_t T = std::min<uint64_t>(std::max(1u, std::min(hw ? hw : 1u, 8u)), n / kMinPart + 1);
    if (T <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const uint64_t per = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (uint64_t t = 1; t < T; ++t) {
        const uint64_t b = std::min(n, t * per), e = std::min(n, b + per);
        th.emplace_back([=] { std::memcpy(dst + b, src + b, e - b); });
    }
    std::memcpy(dst, src, std::min(n, per));
    for (std::thread &x : th) x.join();
}

// Wait for the queued scan and take the site count.
int settle(issl_extractor *ex) {
    if (!ex->in_flight) return ISSL_OK;
    ISSL_HIP(hipStreamSynchronize(ex->stream));
    ex->n_keys = *static_cast<const uint64_t *>(ex->h_count.p);
    ex->in_flight = false;
    return ISSL_OK;
}

// Queue the scan of the first nbytes of pin[cur] (start positions
// [0, npos)) and switch to the other buffer.  The previous scan is settled
// first, so its count sizes the key array exactly and the buffer switched
// to is free again.
int submit(issl_extractor *ex, uint64_t nbytes, uint64_t npos) {
    int rc;
    if ((rc = settle(ex)) != ISSL_OK) return rc;
    const int b = ex->cur;
    ex->cur ^= 1;
    if (npos == 0 || nbytes < kWin) return ISSL_OK;
    if ((rc = ensure_cap(ex, ex->n_keys + 2 * npos)) != ISSL_OK) return rc;
    ISSL_HIP(hipMemcpyAsync(ex->d_seq[b].p, ex->pin[b].p, nbytes, hipMemcpyHostToDevice, ex->stream));
    hipLaunchKernelGGL(k_pam_scan, dim3(grid_for(npos, kScanBlock, 4096)), dim3(kScanBlock), 0,
                       ex->stream, static_cast<const uint8_t *>(ex->d_seq[b].p), nbytes, npos, ex->d_keys,
                       static_cast<unsigned long long *>(ex->d_count.p));
    ISSL_HIP(hipGetLastError());
    ISSL_HIP(hipMemcpyAsync(ex->h_count.p, ex->d_count.p, 8, hipMemcpyDeviceToHost, ex->stream));
    ex->in_flight = true;
    return ISSL_OK;
}

int flush(issl_extractor *ex) {
    const uint64_t n = ex->pending;
    ex->pending = 0;
    return n ? submit(ex, n, n) : ISSL_OK;
}

// Flush and settle: every site added so far is counted.
int drain(issl_extractor *ex) {
    const int rc = flush(ex);
    return rc != ISSL_OK ? rc : settle(ex);
}

int add_sequence(issl_extractor *ex, const char *seq, uint64_t len) {
    ex->finished = false;
    ex->bases += len;
    if (len < kWin) return ISSL_OK;  // holds no window
    int rc;
    if (len <= ex->chunk) {
        // Packed with the others, a separator byte after it.
        if (ex->pending + len + 1 > ex->chunk + kWin && (rc = flush(ex)) != ISSL_OK) return rc;
        char *pin = static_cast<char *>(ex->pin[ex->cur].p);
        par_copy(pin + ex->pending, seq, len);
        pin[ex->pending + len] = '\n';
        ex->pending += len + 1;
        return ISSL_OK;
    }
    if ((rc = flush(ex)) != ISSL_OK) return rc;
    // Pieces of `chunk` start positions, each with the 22 bases after it.
    for (uint64_t off = 0; off < len; off += ex->chunk) {
        const uint64_t npos = std::min<uint64_t>(ex->chunk, len - off);
        const uint64_t nb = std::min<uint64_t>(len - off, ex->chunk + kWin - 1);
        par_copy(static_cast<char *>(ex->pin[ex->cur].p), seq + off, nb);
        if ((rc = submit(ex, nb, npos)) != ISSL_OK) return rc;
    }
    return ISSL_OK;
}

// boost::algorithm::trim's default set (std::isspace in the C locale).
inline bool is_ws(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

void trim(const char *&b, const char *&e) {
    while (b < e && is_ws(*b)) ++b;
    while (e > b && is_ws(e[-1])) --e;
}

// One input file, split as ExtractOfftargets.cpp:76-109 splits it.
int add_file(issl_extractor *ex, const std::string &path) {
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) return ISSL_OK;  // the reference reads nothing from it
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size == 0) {
        close(fd);
        return ISSL_OK;  // one empty sequence: no sites
    }
    const size_t len = static_cast<size_t>(st.st_size);
    void *map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return fail(ISSL_ERR_IO, "cannot map " + path);
    const char *data = static_cast<const char *>(map), *end = data + len;
    int rc = ISSL_OK;
    auto next_line = [&](const char *&p, const char *&lb, const char *&le) {
        if (p >= end) return false;
        lb = p;
        const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
        le = nl ? nl : end;
        p = nl ? nl + 1 : end;
        return true;
    };
    const char *p = data, *lb, *le;
    if (data[0] == '>') {
        // FASTA: the first line is a header; later lines are trimmed, a '>'
        // line starts the next sequence, other lines are appended.
        next_line(p, lb, le);
        std::string cur;
        while (rc == ISSL_OK && next_line(p, lb, le)) {
            trim(lb, le);
            if (lb < le && *lb == '>') {
                rc = add_sequence(ex, cur.data(), cur.size());
                cur.clear();
            } else {
                cur.append(lb, static_cast<size_t>(le - lb));
            }
        }
        if (rc == ISSL_OK) rc = add_sequence(ex, cur.data(), cur.size());
    } else {
        // Plain text: every line (trimmed) is a sequence.
        while (rc == ISSL_OK && next_line(p, lb, le)) {
            trim(lb, le);
            rc = add_sequence(ex, lb, static_cast<uint64_t>(le - lb));
        }
    }
    munmap(map, len);
    return rc;
}

int need_finished(const issl_extractor *ex) {
    if (!ex) return fail(ISSL_ERR_ARG, "null extractor");
    if (!ex->finished) return fail(ISSL_ERR_ARG, "call issl_extractor_finish first");
    return ISSL_OK;
}

}  // namespace

extern "C" int issl_extractor_create(int device, uint64_t chunk_bytes, issl_extractor **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    *out = nullptr;
    const int ndev = issl_device_count();
    if (ndev <= 0) return fail(ISSL_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(ISSL_ERR_ARG, "device ordinal out of range");
    if (chunk_bytes != 0 && chunk_bytes < 64) return fail(ISSL_ERR_ARG, "chunk_bytes must be 0 or at least 64");
    std::unique_ptr<issl_extractor> ex(new (std::nothrow) issl_extractor);
    if (!ex) return fail(ISSL_ERR_NOMEM, "host allocation failed");
    ex->device = device;
    ex->chunk = chunk_bytes ? chunk_bytes : kDefaultChunk;
    DeviceGuard guard(device);
    ISSL_HIP(hipStreamCreateWithFlags(&ex->stream, hipStreamNonBlocking));
    int rc;
    for (int b = 0; b < 2; ++b) {
        if ((rc = ex->pin[b].reserve(ex->chunk + kWin)) != ISSL_OK) return rc;
        if ((rc = ex->d_seq[b].reserve(ex->chunk + kWin)) != ISSL_OK) return rc;
    }
    if ((rc = ex->h_count.reserve(8)) != ISSL_OK) return rc;
    if ((rc = ex->d_count.reserve(8)) != ISSL_OK) return rc;
    ISSL_HIP(hipMemsetAsync(ex->d_count.p, 0, 8, ex->stream));
    *static_cast<uint64_t *>(ex->h_count.p) = 0;
    ISSL_HIP(hipStreamSynchronize(ex->stream));
    *out = ex.release();
    return ok();
}

extern "C" int issl_extractor_destroy(issl_extractor *ex) {
    if (!ex) return ok();
    DeviceGuard guard(ex->device);
    if (ex->stream) (void)hipStreamSynchronize(ex->stream);
    if (ex->d_keys) (void)hipFree(ex->d_keys);
    for (int b = 0; b < 2; ++b) {
        ex->d_seq[b].release();
        ex->pin[b].release();
    }
    ex->h_count.release();
    ex->d_count.release();
    ex->d_sorted.release();
    ex->d_uniq.release();
    ex->d_occ.release();
    ex->d_nruns.release();
    ex->d_tmp.release();
    ex->d_out.release();
    if (ex->stream) (void)hipStreamDestroy(ex->stream);
    delete ex;
    return ok();
}

extern "C" int issl_extractor_add_sequence(issl_extractor *ex, const char *seq, uint64_t len) {
    if (!ex) return fail(ISSL_ERR_ARG, "null extractor");
    if (len && !seq) return fail(ISSL_ERR_ARG, "null sequence");
    DeviceGuard guard(ex->device);
    const int rc = add_sequence(ex, seq, len);
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_extractor_add_path(issl_extractor *ex, const char *path) {
    if (!ex || !path) return fail(ISSL_ERR_ARG, "null argument");
    DeviceGuard guard(ex->device);
    namespace fs = std::filesystem;
    std::error_code ec;
    const fs::path p(path);
    int rc = ISSL_OK;
    if (fs::is_regular_file(p, ec)) {
        rc = add_file(ex, p.string());
    } else if (fs::is_directory(p, ec)) {
        std::vector<std::string> files;
        for (const fs::directory_entry &e : fs::directory_iterator(p, ec))
            if (e.is_regular_file(ec)) files.push_back(e.path().string());
        std::sort(files.begin(), files.end());
        for (const std::string &f : files)
            if ((rc = add_file(ex, f)) != ISSL_OK) break;
    } else {
        // ExtractOfftargets.cpp:55-60
        printf("Error processing: %s\nSkipping entry, Please check the path and try again.\n", path);
        fflush(stdout);
    }
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_extractor_finish(issl_extractor *ex, uint64_t *n_sites, uint64_t *n_distinct) {
    if (!ex) return fail(ISSL_ERR_ARG, "null extractor");
    DeviceGuard guard(ex->device);
    int rc = drain(ex);
    if (rc != ISSL_OK) return rc;
    const uint64_t n = ex->n_keys;
    if (n > kMaxSites)
        return fail(ISSL_ERR_ARG, std::to_string(n) + " sites: at most 2^31 - 1 per finish");
    ex->n_distinct = 0;
    if (n > 0) {
        const int ni = static_cast<int>(n);
        if ((rc = ex->d_sorted.reserve(n * 8)) != ISSL_OK || (rc = ex->d_uniq.reserve(n * 8)) != ISSL_OK ||
            (rc = ex->d_occ.reserve(n * 4)) != ISSL_OK || (rc = ex->d_nruns.reserve(8)) != ISSL_OK)
            return rc;
        uint64_t *sorted = static_cast<uint64_t *>(ex->d_sorted.p);
        uint64_t *uniq = static_cast<uint64_t *>(ex->d_uniq.p);
        uint32_t *occ = static_cast<uint32_t *>(ex->d_occ.p);
        int *nruns = static_cast<int *>(ex->d_nruns.p);
        size_t t_sort = 0, t_rle = 0;
        ISSL_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, ex->d_keys, sorted, ni, 0, kKeyBits,
                                                   ex->stream));
        ISSL_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, t_rle, sorted, uniq, occ, nruns, ni,
                                                       ex->stream));
        if ((rc = ex->d_tmp.reserve(std::max(t_sort, t_rle))) != ISSL_OK) return rc;
        ISSL_HIP(hipcub::DeviceRadixSort::SortKeys(ex->d_tmp.p, t_sort, ex->d_keys, sorted, ni, 0,
                                                   kKeyBits, ex->stream));
        ISSL_HIP(hipcub::DeviceRunLengthEncode::Encode(ex->d_tmp.p, t_rle, sorted, uniq, occ, nruns, ni,
                                                       ex->stream));
        int runs = 0;
        ISSL_HIP(hipMemcpyAsync(&runs, nruns, sizeof(int), hipMemcpyDeviceToHost, ex->stream));
        ISSL_HIP(hipStreamSynchronize(ex->stream));
        ex->n_distinct = static_cast<uint64_t>(runs);
    }
    ex->finished = true;
    if (n_sites) *n_sites = n;
    if (n_distinct) *n_distinct = ex->n_distinct;
    return ok();
}

extern "C" int issl_extractor_sites(issl_extractor *ex, uint64_t *sigs, uint32_t *occ) {
    int rc = need_finished(ex);
    if (rc != ISSL_OK) return rc;
    const uint64_t n = ex->n_distinct;
    if (n == 0) return ok();
    DeviceGuard guard(ex->device);
    if (sigs) {
        if ((rc = ex->d_out.reserve(n * 8)) != ISSL_OK) return rc;
        hipLaunchKernelGGL(k_key_sig, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ex->stream,
                           static_cast<const uint64_t *>(ex->d_uniq.p), n,
                           static_cast<uint64_t *>(ex->d_out.p));
        ISSL_HIP(hipGetLastError());
        ISSL_HIP(hipMemcpyAsync(sigs, ex->d_out.p, n * 8, hipMemcpyDeviceToHost, ex->stream));
    }
    if (occ) ISSL_HIP(hipMemcpyAsync(occ, ex->d_occ.p, n * 4, hipMemcpyDeviceToHost, ex->stream));
    ISSL_HIP(hipStreamSynchronize(ex->stream));
    return ok();
}

extern "C" int issl_extractor_write(issl_extractor *ex, const char *path) {
    int rc = need_finished(ex);
    if (rc != ISSL_OK) return rc;
    if (!path) return fail(ISSL_ERR_ARG, "null path");
    DeviceGuard guard(ex->device);
    FILE *f = fopen(path, "wb");
    if (!f) return fail(ISSL_ERR_IO, std::string("cannot open ") + path + " for writing");
    const uint64_t n = ex->n_keys;
    const uint64_t per = std::min<uint64_t>(std::max<uint64_t>(n, 1), kFormatSites);
    const size_t line = kSiteLen + 1;
    HostBuf host;
    if ((rc = ex->d_out.reserve(per * line)) != ISSL_OK || (rc = host.reserve(per * line)) != ISSL_OK) {
        fclose(f);
        return rc;
    }
    const uint64_t *sorted = static_cast<const uint64_t *>(ex->d_sorted.p);
    for (uint64_t i = 0; i < n && rc == ISSL_OK; i += per) {
        const uint64_t m = std::min<uint64_t>(per, n - i);
        hipLaunchKernelGGL(k_format, dim3(grid_for(m, 256, 8192)), dim3(256), 0, ex->stream, sorted + i, m,
                           static_cast<char *>(ex->d_out.p));
        if (hipMemcpyAsync(host.p, ex->d_out.p, m * line, hipMemcpyDeviceToHost, ex->stream) != hipSuccess ||
            hipStreamSynchronize(ex->stream) != hipSuccess)
            rc = fail(ISSL_ERR_DEVICE, "formatting the sites failed");
        else if (fwrite(host.p, 1, m * line, f) != m * line)
            rc = fail(ISSL_ERR_IO, std::string("write to ") + path + " failed");
    }
    if (fclose(f) != 0 && rc == ISSL_OK) rc = fail(ISSL_ERR_IO, std::string("closing ") + path + " failed");
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_extractor_build_index(issl_extractor *ex, const uint64_t *slice_masks,
                                          uint64_t n_slices, issl_index **out) {
    int rc = need_finished(ex);
    if (rc != ISSL_OK) return rc;
    DeviceGuard guard(ex->device);
    const uint64_t n = ex->n_distinct;
    if ((rc = ex->d_out.reserve(std::max<uint64_t>(n, 1) * 8)) != ISSL_OK) return rc;
    if (n) {
        hipLaunchKernelGGL(k_key_sig, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ex->stream,
                           static_cast<const uint64_t *>(ex->d_uniq.p), n,
                           static_cast<uint64_t *>(ex->d_out.p));
        ISSL_HIP(hipGetLastError());
    }
    ISSL_HIP(hipStreamSynchronize(ex->stream));
    rc = index_build(n, kSiteLen, n_slices, static_cast<const uint64_t *>(ex->d_out.p),
                     static_cast<const uint32_t *>(ex->d_occ.p), true, slice_masks, ex->device, out);
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_extractor_stats(issl_extractor *ex, uint64_t *bases, uint64_t *sites) {
    if (!ex) return fail(ISSL_ERR_ARG, "null extractor");
    DeviceGuard guard(ex->device);
    const int rc = drain(ex);  // packed sequences not yet scanned, the queued scan
    if (rc != ISSL_OK) return rc;
    if (bases) *bases = ex->bases;
    if (sites) *sites = ex->n_keys;
    return ok();
}
