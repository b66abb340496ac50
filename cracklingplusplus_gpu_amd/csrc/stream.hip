// stream.hip -- streamed ISSL index: the posting lists stay in host memory (a
// memory-mapped .issl file or the caller's arrays) and pass through a bounded
// set of device slice buffers, a window of slices at a time.  The reference's
// memory-mapped scorer reads the same file layout through a mapping instead
// of into memory (ISSLScoreOfftargetsMMF/ISSLScoreOfftargetsMMF.cpp:99-153,
// lib/ISSLScoringModuleMMF.cpp:53-150); here the mapping feeds HBM, so an
// index larger than the device memory can be scored.
//
// Resident on the device: signatures (8 B per site), list starts, slice
// metadata and tables, plus `slots` slice buffers of N postings and N
// signature-inline words each (16 B per site per buffer).
//
// One scoring call runs windows of G consecutive slices in ascending slice
// order.  Window k's postings are copied into slot group k % 2 on the copy
// stream while the previous window is scored on the caller's stream.  Every
// guide's sums, counts and exit flag carry from one window to the next
// (GuideState, two copies used alternately so each launch reads one and
// writes the other), and the kernels continue the same sequence of additions:
// the PLAIN form stays bit-exact with the resident index, and the early exit
// stops at the same site.  With S or more slots every slice stays resident
// after its first upload and a call is a single window.
//
// Posting ids are checked against N while they are copied into the pinned
// staging chunks, before any kernel can read them.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "device.hpp"
#include "issl_index.h"

namespace issl {

namespace {

constexpr uint64_t kChunkWords = uint64_t(8) << 20;  // 64 MiB staging chunks
constexpr int kStageChunks = 3;
constexpr uint64_t kSlotBytesPerSite = 16;          // posting + inline word

// Copies one staging chunk with several host threads and checks every
// posting id against N on the way.
class CopyPool {
public:
    explicit CopyPool(int threads) : nthreads_(static_cast<uint64_t>(threads)) {
        for (int t = 0; t < threads; ++t) th_.emplace_back([this, t] { work(t); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }
    CopyPool(const CopyPool &) = delete;
    CopyPool &operator=(const CopyPool &) = delete;

    // Returns the number of entries whose id is >= n_sites.
    uint64_t copy(uint64_t *dst, const uint64_t *src, uint64_t count, uint64_t n_sites) {
        std::unique_lock<std::mutex> lk(mu_);
        dst_ = dst;
        src_ = src;
        count_ = count;
        n_sites_ = n_sites;
        bad_ = 0;
        pending_ = static_cast<int>(nthreads_);
        ++gen_;
        cv_.notify_all();
        done_.wait(lk, [this] { return pending_ == 0; });
        return bad_;
    }

private:
    void work(int t) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t *dst;
            const uint64_t *src;
            uint64_t count, n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                dst = dst_;
                src = src_;
                count = count_;
                n = n_sites_;
            }
            const uint64_t T = nthreads_;
            const uint64_t per = (count + T - 1) / T;
            const uint64_t b = std::min<uint64_t>(count, per * t);
            const uint64_t e = std::min<uint64_t>(count, b + per);
            uint64_t bad = 0;
            for (uint64_t i = b; i < e; ++i) {
                const uint64_t v = src[i];
                dst[i] = v;
                bad += (v & 0xFFFFFFFFULL) >= n;
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                bad_ += bad;
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }

    const uint64_t nthreads_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
    uint64_t *dst_ = nullptr;
    const uint64_t *src_ = nullptr;
    uint64_t count_ = 0, n_sites_ = 0, bad_ = 0;
};

}  // namespace

struct StreamSource {
    std::vector<const uint64_t *> entries;   // [S] host postings of each slice (N each)
    void *map = nullptr;                     // file mapping (null: the caller's arrays)
    size_t map_len = 0;
    int slots = 0;                           // device slice buffers
    uint64_t *d_post = nullptr;              // [slots][N] postings occ<<32 | id
    uint64_t *d_inline = nullptr;            // [slots][N] sig | occ<<40
    std::vector<int> slot_slice;             // slice held by each buffer (-1: none)
    std::vector<char> slot_inline;           // its inline words are built
    hipStream_t copy = nullptr;
    hipEvent_t ready[2] = {nullptr, nullptr};  // a slot group's uploads are done
    hipEvent_t freed[2] = {nullptr, nullptr};  // a slot group's last window is scored
    bool freed_set[2] = {false, false};
    HostBuf stage;                           // kStageChunks pinned chunks
    uint64_t chunk_words = 0;                // entries per chunk
    hipEvent_t stage_ev[kStageChunks] = {};
    bool stage_used[kStageChunks] = {};
    int stage_next = 0;
    std::unique_ptr<CopyPool> pool;
    uint64_t uploaded_bytes = 0;
    uint64_t windows = 0;
};

__global__ void k_inline_slot(const uint64_t *__restrict__ post, const uint64_t *__restrict__ sig,
                              uint64_t *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t e = post[i];
        const uint32_t occ = static_cast<uint32_t>(e >> 32);
        const uint64_t o = occ < kOccEscape ? occ : kOccEscape;
        out[i] = (sig[static_cast<uint32_t>(e)] & kSigBits) | (o << 40);
    }
}

void stream_free(issl_index *ix) {
    StreamSource *s = ix->src;
    if (!s) return;
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    if (s->d_post) (void)hipFree(s->d_post);
    if (s->d_inline) (void)hipFree(s->d_inline);
    for (hipEvent_t e : s->ready)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->freed)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->stage_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    s->stage.release();
    if (s->map) munmap(s->map, s->map_len);
    delete s;
    ix->src = nullptr;
}

namespace {

// Copy slice s's postings into slot j on the copy stream (ids checked first).
int upload_slice(issl_index *ix, StreamSource &src, int s, int j) {
    const uint64_t N = ix->n;
    const uint64_t *h = src.entries[static_cast<size_t>(s)];
    uint64_t *dst = src.d_post + static_cast<uint64_t>(j) * N;
    uint64_t *pin = static_cast<uint64_t *>(src.stage.p);
    src.slot_slice[static_cast<size_t>(j)] = -1;
    src.slot_inline[static_cast<size_t>(j)] = 0;
    for (uint64_t off = 0; off < N; off += src.chunk_words) {
        const uint64_t c = std::min<uint64_t>(src.chunk_words, N - off);
        const int k = src.stage_next;
        src.stage_next = (k + 1) % kStageChunks;
        if (src.stage_used[k]) ISSL_HIP(hipEventSynchronize(src.stage_ev[k]));
        uint64_t *chunk = pin + static_cast<uint64_t>(k) * src.chunk_words;
        if (src.pool->copy(chunk, h + off, c, N) != 0)
            return fail(ISSL_ERR_FORMAT, "slice " + std::to_string(s) + " posting entries hold ids >= N");
        ISSL_HIP(hipMemcpyAsync(dst + off, chunk, c * 8, hipMemcpyHostToDevice, src.copy));
        ISSL_HIP(hipEventRecord(src.stage_ev[k], src.copy));
        src.stage_used[k] = true;
        src.uploaded_bytes += c * 8;
    }
    src.slot_slice[static_cast<size_t>(j)] = s;
    return ISSL_OK;
}

// Signature-inline words of slot j, on the scoring stream (the copy stream
// then moves the next window's postings meanwhile).
int build_slot_inline(issl_index *ix, StreamSource &src, int j, hipStream_t st) {
    const uint64_t N = ix->n;
    if (N == 0) {
        src.slot_inline[static_cast<size_t>(j)] = 1;
        return ISSL_OK;
    }
    const uint64_t off = static_cast<uint64_t>(j) * N;
    hipLaunchKernelGGL(k_inline_slot, dim3(2048), dim3(256), 0, st, src.d_post + off,
                       ix->d_sig, src.d_inline + off, N);
    ISSL_HIP(hipGetLastError());
    src.slot_inline[static_cast<size_t>(j)] = 1;
    return ISSL_OK;
}

}  // namespace

int score_streamed(issl_index *ix, const uint64_t *d_guides, uint64_t n, const issl_params *p,
                   const ScoreScalars &sc, const DevResults &o, uint64_t *distinct,
                   hipStream_t st) {
    StreamSource &src = *ix->src;
    const int S = static_cast<int>(ix->S);
    const bool batched = p->form == ISSL_FORM_BATCHED;  // both forms read the inline words
    // Window size and slot groups: everything resident, or two groups of G
    // slots used alternately (one group when there is a single slot).
    int G, groups;
    if (src.slots >= S) {
        G = std::max(S, 1);
        groups = 1;
    } else if (src.slots >= 2) {
        G = src.slots / 2;
        groups = 2;
    } else {
        G = 1;
        groups = 1;
    }
    const int nwin = S == 0 ? 1 : (S + G - 1) / G;
    int rc;
    GuideState *state[2] = {nullptr, nullptr};
    if (nwin > 1) {
        if ((rc = ix->ws.state.reserve(2 * n * sizeof(GuideState))) != ISSL_OK) return rc;
        state[0] = static_cast<GuideState *>(ix->ws.state.p);
        state[1] = state[0] + n;
    }
    for (int k = 0; k < nwin; ++k) {
        const int grp = k % groups;
        const int s0 = k * G, s1 = std::min(S, s0 + G);
        // The group's buffers are free once its previous window is scored.
        if (src.freed_set[grp]) ISSL_HIP(hipStreamWaitEvent(src.copy, src.freed[grp], 0));
        for (int s = s0; s < s1; ++s) {
            const int j = grp * G + (s - s0);
            if (src.slot_slice[static_cast<size_t>(j)] != s &&
                (rc = upload_slice(ix, src, s, j)) != ISSL_OK)
                return rc;
        }
        ISSL_HIP(hipEventRecord(src.ready[grp], src.copy));
        ISSL_HIP(hipStreamWaitEvent(st, src.ready[grp], 0));
        for (int s = s0; s < s1; ++s) {
            const int j = grp * G + (s - s0);
            if (!src.slot_inline[static_cast<size_t>(j)] &&
                (rc = build_slot_inline(ix, src, j, st)) != ISSL_OK)
                return rc;
        }

        Window w;
        w.s0 = s0;
        w.s1 = s1;
        const uint64_t base = static_cast<uint64_t>(grp) * G * ix->n;
        w.post = src.d_post + base;
        w.inl = src.d_inline + base;
        w.state_in = k > 0 ? state[(k - 1) & 1] : nullptr;
        w.state_out = k + 1 < nwin ? state[k & 1] : nullptr;
        if (batched) {
            rc = launch_batched(ix, d_guides, n, sc, p->precision, p->chunk, ix->tuning, o, distinct,
                                ix->ws, st, &w);
        } else {
            rc = launch_plain(ix, d_guides, n, sc, o, distinct, nullptr, nullptr, st, &w);
        }
        if (rc != ISSL_OK) return rc;
        ISSL_HIP(hipEventRecord(src.freed[grp], st));
        src.freed_set[grp] = true;
        ++src.windows;
    }
    return ISSL_OK;
}

namespace {

// A streamed index over host postings: entries[s] points at slice s's N
// entries and stays valid for the index's lifetime.
int make_streamed(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices, const uint64_t *sigs,
                  const uint64_t *slice_masks, const std::vector<const uint64_t *> &sizes,
                  const std::vector<const uint64_t *> &entries, int device, uint64_t slot_bytes,
                  void *map, size_t map_len, issl_index **out) {
    // The index's own consistency first, then the device.
    int rc = ISSL_OK;
    if (seq_len == 20 && n_sites < (1ULL << 32) && count_wide_sigs(sigs, n_sites) != 0)
        rc = fail(ISSL_ERR_FORMAT, "signatures hold codes beyond position 19");
    if (rc == ISSL_OK) rc = check_index_args(n_sites, seq_len, n_slices, device);
    if (rc != ISSL_OK) {
        if (map) munmap(map, map_len);
        return rc;
    }
    std::unique_ptr<issl_index, void (*)(issl_index *)> ix(new (std::nothrow) issl_index, free_index);
    if (!ix) {
        if (map) munmap(map, map_len);
        return fail(ISSL_ERR_NOMEM, "host allocation failed");
    }
    ix->device = device;
    ix->src = new (std::nothrow) StreamSource;
    if (!ix->src) {
        if (map) munmap(map, map_len);
        return fail(ISSL_ERR_NOMEM, "host allocation failed");
    }
    StreamSource &src = *ix->src;
    src.map = map;  // released with the index from here on
    src.map_len = map_len;
    src.entries = entries;

    uint64_t lstart_count = 0;
    if ((rc = init_meta(ix.get(), n_sites, seq_len, n_slices, slice_masks, &lstart_count)) != ISSL_OK)
        return rc;
    std::vector<uint32_t> lstart(lstart_count);
    if ((rc = build_lstart(ix.get(), sizes, lstart)) != ISSL_OK) return rc;

    DeviceGuard guard(device);
    // Slice buffers: as many as the budget holds (at most S); a budget of 0
    // takes as many as fit in the free device memory less 4 GiB.
    const uint64_t per_slot = std::max<uint64_t>(n_sites, 1) * kSlotBytesPerSite;
    uint64_t budget = slot_bytes;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        ISSL_HIP(hipMemGetInfo(&fr, &tot));
        const uint64_t keep = (uint64_t(4) << 30) + n_sites * 8;
        budget = fr > keep ? fr - keep : 0;
    }
    const uint64_t want = std::max<uint64_t>(n_slices, 1);
    const uint64_t slots = std::min<uint64_t>(want, budget / per_slot);
    if (slots == 0)
        return fail(ISSL_ERR_NOMEM, "device budget of " + std::to_string(budget) +
                                        " bytes holds no slice buffer (" + std::to_string(per_slot) +
                                        " bytes each)");
    src.slots = static_cast<int>(slots);
    src.slot_slice.assign(slots, -1);
    src.slot_inline.assign(slots, 0);

    if ((rc = alloc_index(ix.get(), lstart_count, false)) != ISSL_OK) return rc;
    const size_t sb = static_cast<size_t>(slots * std::max<uint64_t>(n_sites, 1) * 8);
    if (hipMalloc(&src.d_post, sb) != hipSuccess || hipMalloc(&src.d_inline, sb) != hipSuccess)
        return fail(ISSL_ERR_NOMEM, "device allocation of the slice buffers (" +
                                        std::to_string(2 * sb) + " bytes) failed");
    ix->device_bytes += 2 * sb;
    ISSL_HIP(hipStreamCreateWithFlags(&src.copy, hipStreamNonBlocking));
    for (int g = 0; g < 2; ++g) {
        ISSL_HIP(hipEventCreateWithFlags(&src.ready[g], hipEventDisableTiming));
        ISSL_HIP(hipEventCreateWithFlags(&src.freed[g], hipEventDisableTiming));
    }
    for (int k = 0; k < kStageChunks; ++k)
        ISSL_HIP(hipEventCreateWithFlags(&src.stage_ev[k], hipEventDisableTiming));
    src.chunk_words = std::min<uint64_t>(kChunkWords, std::max<uint64_t>(n_sites, 1));
    if ((rc = src.stage.reserve(kStageChunks * src.chunk_words * 8)) != ISSL_OK) return rc;
    const unsigned hw = std::thread::hardware_concurrency();
    src.pool.reset(new CopyPool(static_cast<int>(std::max(1u, std::min(hw ? hw : 1u, 8u)))));

    if ((rc = upload(ix.get(), ix->d_sig, sigs, n_sites * 8)) != ISSL_OK) return rc;
    if ((rc = upload(ix.get(), ix->d_lstart, lstart.data(), lstart.size() * 4)) != ISSL_OK) return rc;
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    ix->ws.pinned.release();
    *out = ix.release();
    return ISSL_OK;
}

}  // namespace

}  // namespace issl

using namespace issl;

extern "C" int issl_index_from_arrays_streamed(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                                               const uint64_t *sigs, const uint64_t *slice_masks,
                                               const uint64_t *list_sizes, const uint64_t *entries,
                                               int device, uint64_t slot_bytes, issl_index **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    *out = nullptr;
    if ((n_sites && !sigs) || (n_slices && (!slice_masks || !list_sizes)) ||
        (n_sites && n_slices && !entries))
        return fail(ISSL_ERR_ARG, "null index array");
    if (n_slices > static_cast<uint64_t>(kMaxSlices))
        return fail(ISSL_ERR_ARG, "at most 64 slices are supported");
    std::vector<const uint64_t *> sizes(n_slices), ent(n_slices);
    uint64_t base = 0;
    for (uint64_t s = 0; s < n_slices; ++s) {
        uint8_t pos[64];
        const int w = mask_positions(slice_masks[s], seq_len, pos);
        if (w > kMaxWidth) return fail(ISSL_ERR_FORMAT, "slice wider than 16 positions");
        sizes[s] = list_sizes + base;
        base += 1ULL << (2 * w);
        ent[s] = entries + s * n_sites;
    }
    const int rc = make_streamed(n_sites, seq_len, n_slices, sigs, slice_masks, sizes, ent, device,
                                 slot_bytes, nullptr, 0, out);
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_index_open_streamed(const char *path, int device, uint64_t slot_bytes,
                                        issl_index **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    *out = nullptr;
    uint64_t n, L, S, nl;
    int rc = issl_read_header(path, &n, &L, &S, &nl);  // also checks the file size
    if (rc != ISSL_OK) return rc;
    if (S > static_cast<uint64_t>(kMaxSlices)) return fail(ISSL_ERR_ARG, "at most 64 slices are supported");
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(ISSL_ERR_IO, std::string("Error reading index: could not open file ") + path);
    struct stat stt;
    if (fstat(fd, &stt) != 0) {
        close(fd);
        return fail(ISSL_ERR_IO, "stat failed");
    }
    const size_t len = static_cast<size_t>(stt.st_size);
    void *map = mmap(nullptr, len, PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return fail(ISSL_ERR_IO, std::string("cannot map ") + path);
    // Layout: header (N, length, S), N signatures, S masks, then per slice
    // its 4^w list sizes followed by its N entries.
    const uint64_t *w = static_cast<const uint64_t *>(map);
    const uint64_t *sigs = w + 3;
    const uint64_t *masks = sigs + n;
    std::vector<const uint64_t *> sizes(S), ent(S);
    const uint64_t *cur = masks + S;
    for (uint64_t s = 0; s < S; ++s) {
        uint8_t pos[64];
        const int wd = mask_positions(masks[s], L, pos);
        sizes[s] = cur;
        cur += 1ULL << (2 * wd);
        ent[s] = cur;
        cur += n;
    }
    rc = make_streamed(n, L, S, sigs, masks, sizes, ent, device, slot_bytes, map, len, out);
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_index_stream_stats(issl_index *ix, uint64_t *slots, uint64_t *uploaded_bytes,
                                       uint64_t *windows) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    const StreamSource *s = ix->src;
    if (slots) *slots = s ? static_cast<uint64_t>(s->slots) : 0;
    if (uploaded_bytes) *uploaded_bytes = s ? s->uploaded_bytes : 0;
    if (windows) *windows = s ? s->windows : 0;
    return ok();
}
