// device.hip -- resident ISSL index on one MI355X (gfx950), the PLAIN scoring
// kernel and the C ABI of issl_gpu.h.
//
// Device layout of an index (all in HBM, resident until issl_index_close):
//   d_sig    u64[N]        signatures, id order (the .issl sig block)
//   d_post   u64[S*N]      posting entries occ<<32 | id, the .issl order
//   d_inline u64[S*N]      same order, sig | occ<<40 (BATCHED form)
//   d_res    u32[S*N]      same order, list-residual words (the tile kernel's
//                          stream; built when every slice has >= 8 positions)
//   d_lstart u32[sum(4^w+1)] per slice, start of each key's list within the
//                          slice's N entries, plus an end sentinel
// A guide's candidate sequence is the concatenation, slice by slice, of the
// lists its keys select: exactly the order of the reference's nested loops
// (lib/ISSLScoringModule.cpp:231-250).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "device.hpp"
#include "issl_index.h"
#include "issl_server.h"

namespace issl {

// ------------------------------------------------------------ buffers

int DevBuf::reserve(size_t want) {
    if (want <= bytes) return ISSL_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    const size_t sz = std::max<size_t>(want, 4096);
    if (hipMalloc(&p, sz) != hipSuccess) {
        p = nullptr;
        return fail(ISSL_ERR_NOMEM, "device allocation of " + std::to_string(sz) + " bytes failed");
    }
    bytes = sz;
    return ISSL_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int HostBuf::reserve(size_t want) {
    if (want <= bytes) return ISSL_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    const size_t sz = std::max<size_t>(want, 4096);
    if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return fail(ISSL_ERR_NOMEM, "pinned host allocation failed");
    }
    bytes = sz;
    return ISSL_OK;
}

void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

// ------------------------------------------------------------ device helpers

__device__ __forceinline__ uint64_t dev_fold(uint64_t x) {
    return ((x >> 1) & kFoldLo) | (x & kFoldLo);
}

// Combinadic rank of the mismatch position set of a fold mask with 1..4 bits.
__device__ __forceinline__ uint32_t mit_rank(uint64_t f, int k) {
    uint32_t r = (k == 1) ? 0u : (k == 2) ? 20u : (k == 3) ? 210u : 1350u;
#pragma unroll
    for (int i = 1; i <= 4; ++i) {
        if (i > k) break;
        const uint32_t p = static_cast<uint32_t>(__builtin_ctzll(f)) >> 1;
        f &= f - 1;
        uint32_t c;
        if (i == 1)
            c = p;
        else if (i == 2)
            c = p * (p - 1) / 2;
        else if (i == 3)
            c = p * (p - 1) * (p - 2) / 6;
        else
            c = p * (p - 1) * (p - 2) * (p - 3) / 24;
        r += c;
    }
    return r;
}

// CFD factor of one site in FP64 (lib/ISSLScoringModule.cpp:308-361): the
// NGG PAM penalty times the position penalty of every differing position,
// in ascending position order.  Exact matches score 1.
__device__ __forceinline__ double cfd_site64(uint64_t q, uint64_t site, uint64_t f, int dist,
                                             const double *__restrict__ cfd) {
    if (dist == 0) return 1.0;
    double c = cfd[320 + 10];
    while (f) {
        const uint32_t p = static_cast<uint32_t>(__builtin_ctzll(f)) >> 1;
        f &= f - 1;
        const uint32_t g = static_cast<uint32_t>(q >> (2 * p)) & 3u;
        const uint32_t o = static_cast<uint32_t>(site >> (2 * p)) & 3u;
        c *= cfd[(p << 4) | (g << 2) | (o ^ 3u)];
    }
    return c;
}

// Exit tests of lib/ISSLScoringModule.cpp:365-395.
__device__ __forceinline__ bool exit_test(int method, double m, double c, double cut) {
    switch (method) {
    case ISSL_METHOD_MIT_AND_CFD: return m > cut && c > cut;
    case ISSL_METHOD_MIT_OR_CFD: return m > cut || c > cut;
    case ISSL_METHOD_AVG_MIT_CFD: return ((m + c) / 2.0) > cut;
    case ISSL_METHOD_MIT: return m > cut;
    default: return c > cut;
    }
}

// Transform and call (lib/ISSLScoringModule.cpp:404-455).
__device__ void write_result(const DevResults &o, uint64_t g, const ScoreScalars &sc, double smit,
                             double scfd, uint64_t occ_sum, uint64_t site_sum) {
    const double vm = 10000.0 / (100.0 + smit);
    const double vc = 10000.0 / (100.0 + scfd);
    const double t = sc.threshold;
    bool reject;
    switch (sc.method) {
    case ISSL_METHOD_MIT_AND_CFD: reject = vm < t && vc < t; break;
    case ISSL_METHOD_MIT_OR_CFD: reject = vm < t || vc < t; break;
    case ISSL_METHOD_AVG_MIT_CFD: reject = (vm + vc) / 2 < t; break;
    case ISSL_METHOD_MIT: reject = vm < t; break;
    default: reject = vc < t; break;
    }
    if (o.mit) o.mit[g] = sc.calc_mit ? vm : -1.0;
    if (o.cfd) o.cfd[g] = sc.calc_cfd ? vc : -1.0;
    if (o.mit_sum) o.mit_sum[g] = smit;
    if (o.cfd_sum) o.cfd_sum[g] = scfd;
    if (o.occ) o.occ[g] = occ_sum;
    if (o.sites) o.sites[g] = site_sum;
    if (o.pass) o.pass[g] = reject ? '0' : '1';
}

// ------------------------------------------------------------ PLAIN kernel

struct PlainArgs {
    const uint64_t *post;
    const uint64_t *inl;            // signature-inline site words, same order as post
    const uint32_t *lstart;
    const SliceMeta *meta;
    const double *mit64;
    const double *cfd64;
    const uint64_t *guides;
    uint64_t n_sites;
    int n_slices;
    ScoreScalars sc;
    DevResults out;
    uint64_t *distinct;
    const uint64_t *cand_off;
    uint32_t *cand_ids;
    int s0, s1;                     // slice window; postings of slice s at (s - s0) * N
    const GuideState *state_in;     // carried state (null: start from zero)
    GuideState *state_out;          // null: write the results
};

constexpr int kPlainBlock = 256;
constexpr int kPlainPer = 4;  // postings per thread per chunk

// One workgroup per guide.  Each chunk of BLOCK*PER postings of the guide's
// slice-major sequence is read with coalesced loads (consecutive threads,
// consecutive entries) of the signature-inline words, which carry each
// site's signature and occurrences (the posting entry is read only for an
// occurrence escape or a site id), and sites within
// 4 mismatches that no earlier slice lists (the reference's first-seen
// toggle, ISSLScoringModule.cpp:252-258, as a stateless rule: slice s'
// lists the site iff the fold misses all of s'’s positions) become hits.
// Hits are compacted into LDS in sequence order with wave ballots; their
// FP64 terms are formed in parallel and one lane adds them in order and
// applies the early exit after each, so sums, counts and the exit point are
// the reference's.
template <int BLOCK, int PER>
__global__ __launch_bounds__(BLOCK) void k_plain(PlainArgs a) {
    constexpr int WAVES = BLOCK / 64;
    constexpr int CH = BLOCK * PER;
    __shared__ uint64_t s_beg[kMaxSlices];
    __shared__ uint64_t s_pre[kMaxSlices + 1];
    __shared__ uint64_t s_fold[kMaxSlices];
    __shared__ uint32_t s_wcnt[2][PER][WAVES];
    __shared__ double s_hmit[CH];
    __shared__ double s_hcfd[CH];
    __shared__ uint32_t s_hocc[CH];
    __shared__ uint32_t s_accept;
    __shared__ int s_exited;
    __shared__ double s_cfd[kCfdTable];

    const uint64_t g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int i = tid; i < kCfdTable; i += BLOCK) s_cfd[i] = a.cfd64[i];
    // Signatures are 40-bit (20 positions); higher bits are ignored.
    const uint64_t q = a.guides[g] & kSigBits;
    const int S = a.n_slices;

    // Sums and counts so far (earlier slice windows of a streamed index); the
    // additions below continue the same sequence, so the order is unchanged.
    double smit = 0.0, scfd = 0.0;  // meaningful in thread 0
    uint64_t occ_sum = 0, site_sum = 0;
    bool was_exited = false;
    if (a.state_in) {
        const GuideState st = a.state_in[g];
        smit = st.m;
        scfd = st.c;
        occ_sum = st.occ;
        site_sum = st.sites & ~kExited;
        was_exited = (st.sites & kExited) != 0;
    }

    if (tid < S) {
        const uint64_t mf = a.meta[tid].fold;
        const uint32_t lb = a.meta[tid].lbase;
        const uint32_t key = slice_key(q, mf);
        const uint32_t b = a.lstart[lb + key];
        const uint32_t e = a.lstart[lb + key + 1];
        const bool in = tid >= a.s0 && tid < a.s1 && !was_exited;
        s_beg[tid] = in ? static_cast<uint64_t>(tid - a.s0) * a.n_sites + b : 0;
        s_pre[tid + 1] = in ? e - b : 0;
        s_fold[tid] = mf;
    }
    __syncthreads();
    if (tid == 0) {
        s_pre[0] = 0;
        for (int s = 0; s < S; ++s) s_pre[s + 1] += s_pre[s];
        s_exited = 0;
    }
    __syncthreads();
    const uint64_t P = s_pre[S];

    uint64_t emitted = 0;
    const uint64_t cand_base = a.cand_ids ? a.cand_off[g] : 0;
    int cur = 0;
    int parity = 0;

    for (uint64_t base = 0; base < P; base += CH) {
        uint64_t wv[PER], addr[PER], site[PER];
        int sl[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint64_t i = base + static_cast<uint64_t>(k) * BLOCK + tid;
            sl[k] = -1;
            wv[k] = 0;
            addr[k] = 0;
            if (i < P) {
                while (i >= s_pre[cur + 1]) ++cur;
                sl[k] = cur;
                addr[k] = s_beg[cur] + (i - s_pre[cur]);
                wv[k] = a.inl[addr[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) site[k] = wv[k] & kSigBits;

        bool hit[PER];
        uint64_t fold[PER];
        uint32_t rank[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            hit[k] = false;
            fold[k] = 0;
            if (sl[k] >= 0) {
                const uint64_t f = dev_fold(q ^ site[k]);
                fold[k] = f;
                if (__popcll(f) <= 4) {
                    bool first = true;
                    for (int t = 0; t < sl[k]; ++t)
                        if ((f & s_fold[t]) == 0) {
                            first = false;
                            break;
                        }
                    hit[k] = first;
                }
            }
            const uint64_t b = __ballot(hit[k]);
            rank[k] = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
            if (lane == 0) s_wcnt[parity][k][wave] = static_cast<uint32_t>(__popcll(b));
        }
        __syncthreads();
        uint32_t H = 0;
        uint32_t off[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            uint32_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = s_wcnt[parity][k][w];
                before += (w < wave) ? c : 0u;
                all += c;
            }
            off[k] = H + before + rank[k];
            H += all;
        }
        parity ^= 1;
        if (H == 0) continue;  // uniform across the workgroup

#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!hit[k]) continue;
            const int dist = __popcll(fold[k]);
            uint32_t occ = static_cast<uint32_t>(wv[k] >> 40);
            if (occ == kOccEscape) occ = static_cast<uint32_t>(a.post[addr[k]] >> 32);
            double tm = 0.0, tc = 0.0;
            if (a.sc.calc_mit && dist > 0) tm = a.mit64[mit_rank(fold[k], dist)] * static_cast<double>(occ);
            if (a.sc.calc_cfd) tc = cfd_site64(q, site[k], fold[k], dist, s_cfd) * static_cast<double>(occ);
            s_hmit[off[k]] = tm;
            s_hcfd[off[k]] = tc;
            s_hocc[off[k]] = occ;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = H;
            int ex = 0;
            for (uint32_t h = 0; h < H; ++h) {
                smit += s_hmit[h];
                scfd += s_hcfd[h];
                occ_sum += s_hocc[h];
                site_sum += 1;
                if (a.sc.use_cut && exit_test(a.sc.method, smit, scfd, a.sc.cut)) {
                    acc = h + 1;
                    ex = 1;
                    break;
                }
            }
            s_accept = acc;
            s_exited = ex;
        }
        __syncthreads();
        const uint32_t acc = s_accept;
        if (a.distinct || a.cand_ids) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (!hit[k] || off[k] >= acc) continue;
                const uint32_t id = static_cast<uint32_t>(a.post[addr[k]]);
                if (a.distinct)
                    atomicOr(reinterpret_cast<unsigned long long *>(a.distinct) + (id >> 6),
                             1ULL << (id & 63));
                if (a.cand_ids) a.cand_ids[cand_base + emitted + off[k]] = id;
            }
        }
        emitted += acc;
        if (s_exited) break;
    }
    if (tid != 0) return;
    if (a.state_out) {
        GuideState st;
        st.m = smit;
        st.c = scfd;
        st.occ = occ_sum;
        st.sites = site_sum | ((was_exited || s_exited) ? kExited : 0);
        a.state_out[g] = st;
    } else {
        write_result(a.out, g, a.sc, smit, scfd, occ_sum, site_sum);
    }
}

int launch_plain(issl_index *ix, const uint64_t *d_guides, uint64_t n, const ScoreScalars &sc,
                 const DevResults &out, uint64_t *distinct, const uint64_t *cand_off,
                 uint32_t *cand_ids, hipStream_t stream, const Window *win) {
    if (n == 0) return ISSL_OK;
    if (n > 0x7FFFFFFFULL) return fail(ISSL_ERR_ARG, "too many guides for one launch");
    if (win && cand_ids) return fail(ISSL_ERR_ARG, "candidate lists need a resident index");
    if (!win && !ix->d_post) return fail(ISSL_ERR_ARG, "this operation needs a resident index");
    PlainArgs a;
    a.post = win ? win->post : ix->d_post;
    a.inl = win ? win->inl : ix->d_inline;
    a.s0 = win ? win->s0 : 0;
    a.s1 = win ? win->s1 : static_cast<int>(ix->S);
    a.state_in = win ? win->state_in : nullptr;
    a.state_out = win ? win->state_out : nullptr;
    a.lstart = ix->d_lstart;
    a.meta = ix->d_meta;
    a.mit64 = ix->d_mit64;
    a.cfd64 = ix->d_cfd64;
    a.guides = d_guides;
    a.n_sites = ix->n;
    a.n_slices = static_cast<int>(ix->S);
    a.sc = sc;
    a.out = out;
    a.distinct = distinct;
    a.cand_off = cand_off;
    a.cand_ids = cand_ids;
    const int rc = timed_dispatch(ix, stream, [&](hipEvent_t e0, hipEvent_t e1) {
        hipExtLaunchKernelGGL((k_plain<kPlainBlock, kPlainPer>), dim3(static_cast<uint32_t>(n)),
                              dim3(kPlainBlock), 0, stream, e0, e1, 0, a);
        return ISSL_OK;
    });
    if (rc != ISSL_OK) return rc;
    ISSL_HIP(hipGetLastError());
    return ISSL_OK;
}

// ------------------------------------------------------------ index kernels

// Signature-inline site words: sig | min(occ, escape)<<40, same order as the
// postings.
__global__ void k_inline(const uint64_t *__restrict__ post, const uint64_t *__restrict__ sig,
                         uint64_t *__restrict__ out, uint64_t total) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t e = post[i];
        const uint32_t occ = static_cast<uint32_t>(e >> 32);
        const uint64_t o = occ < kOccEscape ? occ : kOccEscape;
        out[i] = (sig[static_cast<uint32_t>(e)] & kSigBits) | (o << 40);
    }
}

// List-residual words of one slice: the codes at the positions outside the
// slice's fold mask, ascending, then min(occ, escape) from bit `shift`.
__global__ void k_residual(const uint64_t *__restrict__ post, const uint64_t *__restrict__ sig,
                           uint32_t *__restrict__ out, uint64_t n, uint64_t fold, uint32_t shift) {
    const uint32_t esc = 0xFFFFFFFFu >> shift;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t e = post[i];
        const uint32_t occ = static_cast<uint32_t>(e >> 32);
        out[i] = residual_code(sig[static_cast<uint32_t>(e)], fold) | (std::min(occ, esc) << shift);
    }
}

// Index check: each entry of slice s sits in the list of its own key, its id
// is < N, and every id occurs exactly once in the slice.
__global__ void k_validate(const uint64_t *__restrict__ post, const uint64_t *__restrict__ sig,
                           const uint32_t *__restrict__ lstart, SliceMeta m, uint64_t s,
                           uint64_t n, uint32_t *__restrict__ seen,
                           unsigned long long *__restrict__ bad) {
    const uint32_t nkeys = 1u << (2 * m.width);
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        // Largest key whose list starts at or before i.
        uint32_t lo = 0, hi = nkeys;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (lstart[m.lbase + mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        const uint64_t e = post[s * n + i];
        const uint64_t id = e & 0xFFFFFFFFULL;
        bool good = id < n;
        if (good) {
            good = slice_key(sig[id] & kSigBits, m.fold) == lo;
            if (atomicAdd(seen + id, 1u) != 0u) good = false;
        }
        if (!good) atomicAdd(bad, 1ULL);
    }
}

void timing_events(issl_index *ix, hipEvent_t *start, hipEvent_t *stop) {
    *start = *stop = nullptr;
    if (!ix->timing) return;
    while (ix->ev_pool.size() < ix->ev_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        ix->ev_pool.push_back(e);
    }
    *start = ix->ev_pool[ix->ev_used++];
    *stop = ix->ev_pool[ix->ev_used++];
}

int build_inline_layout(issl_index *ix) {
    const uint64_t total = ix->S * ix->n;
    if (total == 0) return ISSL_OK;
    hipLaunchKernelGGL(k_inline, dim3(4096), dim3(256), 0, ix->stream, ix->d_post, ix->d_sig,
                       ix->d_inline, total);
    if (ix->d_res)
        for (uint64_t s = 0; s < ix->S; ++s)
            hipLaunchKernelGGL(k_residual, dim3(2048), dim3(256), 0, ix->stream, ix->d_post + s * ix->n,
                               ix->d_sig, ix->d_res + s * ix->n, ix->n, ix->meta_h[s].fold,
                               ix->res_shift);
    ISSL_HIP(hipGetLastError());
    return ISSL_OK;
}

// Whether the list-residual words can be built: every slice within positions
// 0..19 and at least kResMinWidth wide; *shift = 2 (20 - narrowest width).
bool residual_layout_fits(const issl_index *ix, uint32_t *shift) {
    if (ix->S == 0) return false;
    uint32_t min_w = kResPositions;
    for (const SliceMeta &m : ix->meta_h) {
        for (uint32_t j = 0; j < m.width; ++j)
            if (m.pos[j] >= kResPositions) return false;
        min_w = std::min(min_w, m.width);
    }
    if (min_w < static_cast<uint32_t>(kResMinWidth)) return false;
    *shift = 2 * (kResPositions - min_w);
    return true;
}

}  // namespace issl

using namespace issl;

// ============================================================ C ABI

namespace issl {

void free_index(issl_index *ix) {
    if (!ix) return;
    (void)hipSetDevice(ix->device);
    if (ix->stream) (void)hipStreamSynchronize(ix->stream);
    // Bound streams may still run calls that use their workspaces (and the
    // caller may have destroyed them): wait for the device.
    if (!ix->bound.empty()) (void)hipDeviceSynchronize();
    ix->bound.clear();
    stream_free(ix);
    void *ptrs[] = {ix->d_sig, ix->d_post, ix->d_inline, ix->d_res, ix->d_lstart, ix->d_meta,
                    ix->d_mit64, ix->d_mit32, ix->d_mitf, ix->d_cfd64, ix->d_cfd32, ix->d_distinct};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ix->ws.guides.release();
    ix->ws.out.release();
    ix->ws.tiles.release();
    ix->ws.scratch.release();
    ix->ws.state.release();
    ix->ws.pinned.release();
    for (hipEvent_t e : ix->ev_pool) (void)hipEventDestroy(e);
    for (GraphEntry &gr : ix->graphs) (void)hipGraphExecDestroy(gr.exec);
    if (ix->ws_done) (void)hipEventDestroy(ix->ws_done);
    if (ix->stream) (void)hipStreamDestroy(ix->stream);
    delete ix;
}

template <typename T>
int dev_alloc(issl_index *ix, T **p, uint64_t count) {
    const size_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    void *q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess)
        return fail(ISSL_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    *p = static_cast<T *>(q);
    ix->device_bytes += bytes;
    return ISSL_OK;
}

// Host -> device copy of a large pageable array through pinned staging.
int upload(issl_index *ix, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return ISSL_OK;
    const size_t stage = size_t(64) << 20;
    if (bytes <= stage) {
        ISSL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return ISSL_OK;
    }
    int rc = ix->ws.pinned.reserve(2 * stage);
    if (rc != ISSL_OK) {
        ISSL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return ISSL_OK;
    }
    char *pin = static_cast<char *>(ix->ws.pinned.p);
    hipEvent_t ev[2];
    ISSL_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    ISSL_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool used[2] = {false, false};
    size_t off = 0;
    int slot = 0;
    int status = ISSL_OK;
    while (off < bytes && status == ISSL_OK) {
        const size_t c = std::min(stage, bytes - off);
        if (used[slot] && hipEventSynchronize(ev[slot]) != hipSuccess) status = ISSL_ERR_DEVICE;
        std::memcpy(pin + slot * stage, static_cast<const char *>(src) + off, c);
        if (hipMemcpyAsync(static_cast<char *>(dst) + off, pin + slot * stage, c,
                           hipMemcpyHostToDevice, ix->stream) != hipSuccess ||
            hipEventRecord(ev[slot], ix->stream) != hipSuccess)
            status = ISSL_ERR_DEVICE;
        used[slot] = true;
        off += c;
        slot ^= 1;
    }
    (void)hipStreamSynchronize(ix->stream);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (status != ISSL_OK) return fail(status, "index upload failed");
    return ISSL_OK;
}

int check_params(const issl_index *ix, const issl_params *p, ScoreScalars *sc) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    if (!p) return fail(ISSL_ERR_ARG, "null params");
    if (p->max_dist != 4)
        return fail(ISSL_ERR_ARG,
                    "max_dist must be 4: the reference scores dist <= 4 regardless of its argument");
    if (p->method < 0 || p->method > 4) return fail(ISSL_ERR_ARG, "Scoring method is not valid.");
    if (p->form != ISSL_FORM_PLAIN && p->form != ISSL_FORM_BATCHED && p->form != ISSL_FORM_ORDERED)
        return fail(ISSL_ERR_ARG, "unknown form");
    if (p->precision != ISSL_PREC_FP64 && p->precision != ISSL_PREC_FP32C)
        return fail(ISSL_ERR_ARG, "unknown precision");
    if (p->form != ISSL_FORM_BATCHED && p->precision != ISSL_PREC_FP64)
        return fail(ISSL_ERR_ARG, "the PLAIN and ORDERED forms compute in FP64 only");
    sc->method = p->method;
    sc->calc_mit = p->method != ISSL_METHOD_CFD;
    sc->calc_cfd = p->method != ISSL_METHOD_MIT;
    sc->threshold = p->threshold;
    // maximum_sum of ISSLScoringModule.cpp:225, formed in host FP64.
    sc->cut = (10000.0 - p->threshold * 100) / p->threshold;
    // t = 0 gives +inf and never exits; the GPU skips the tests then.
    sc->use_cut = !(sc->cut == HUGE_VAL) && !(sc->cut != sc->cut);
    return ISSL_OK;
}

}  // namespace issl

extern "C" int issl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" const char *issl_version(void) { return "issl-mi355x 0.1.0 (gfx950)"; }

extern "C" int issl_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(ISSL_ERR_ARG, "null argument");
    DeviceGuard guard(device);
    size_t fr = 0, tot = 0;
    ISSL_HIP(hipMemGetInfo(&fr, &tot));
    *free_bytes = fr;
    *total_bytes = tot;
    return ok();
}

extern "C" uint64_t issl_index_bytes_needed(uint64_t n_sites, uint64_t n_slices, uint64_t n_lists) {
    // signatures, postings, signature-inline words, list-residual words (not
    // built when a slice is narrower than 8 positions), list starts, slice
    // meta, MIT (f64 + f32), CFD (f64 + f32), distinct bitmap
    return n_sites * 8 + 2 * n_slices * n_sites * 8 + n_slices * n_sites * 4 +
           (n_lists + n_slices) * 4 +
           n_slices * sizeof(SliceMeta) + kMitEntries * 12 + kCfdTable * 12 +
           (n_sites / 64 + 1) * 8;
}

extern "C" int issl_encode(const char *text, uint64_t n, uint64_t stride, uint64_t *out) {
    if ((n && (!text || !out)) || stride < 20) return fail(ISSL_ERR_ARG, "issl_encode: bad arguments");
    for (uint64_t i = 0; i < n; ++i) {
        const char *p = text + i * stride;
        uint64_t s = 0;
        for (unsigned j = 0; j < 20; ++j) {
            const char c = p[j];
            const uint64_t v = c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0;
            s |= v << (2 * j);
        }
        out[i] = s;
    }
    return ok();
}

namespace issl {

// Slice metadata from the masks (ISSLScoringModule.cpp:100-198): positions,
// fold masks, widths and each slice's first list start; *lstart_count is the
// number of list starts (4^w + 1 per slice, with an end sentinel).
int init_meta(issl_index *ix, uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
              const uint64_t *slice_masks, uint64_t *lstart_count) {
    ix->n = n_sites;
    ix->seq_len = seq_len;
    ix->S = n_slices;
    ix->masks.assign(slice_masks, slice_masks + n_slices);
    uint64_t base = 0, lists = 0;
    for (uint64_t s = 0; s < n_slices; ++s) {
        uint8_t pos[64];
        const int w = mask_positions(slice_masks[s], seq_len, pos);
        if (w > kMaxWidth) return fail(ISSL_ERR_FORMAT, "slice wider than 16 positions");
        SliceMeta m{};
        m.width = static_cast<uint32_t>(w);
        m.lbase = static_cast<uint32_t>(base);
        m.fold = 0;
        for (int j = 0; j < w; ++j) {
            m.pos[j] = pos[j];
            m.fold |= 1ULL << (2 * pos[j]);
        }
        for (int p = 0; p < kResPositions; ++p)
            if (!((m.fold >> (2 * p)) & 1ULL)) m.keep |= 1u << p;
        const uint64_t nk = 1ULL << (2 * w);
        if (base + nk + 1 > 0xFFFFFFFFULL)
            return fail(ISSL_ERR_FORMAT, "too many lists for 32-bit list offsets");
        base += nk + 1;
        lists += nk;
        ix->max_w = std::max(ix->max_w, w);
        ix->meta_h.push_back(m);
    }
    ix->n_lists = lists;
    *lstart_count = base;
    return ISSL_OK;
}

// Device arrays of an index (contents uploaded or built by the caller), the
// score tables, the slice metadata and a cleared distinct-site bitmap.
int alloc_index(issl_index *ix, uint64_t lstart_count, bool postings) {
    const uint64_t n_sites = ix->n, n_slices = ix->S;
    ISSL_HIP(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking));
    int rc;
    if ((rc = dev_alloc(ix, &ix->d_sig, n_sites)) != ISSL_OK) return rc;
    if (postings) {
        if ((rc = dev_alloc(ix, &ix->d_post, n_slices * n_sites)) != ISSL_OK) return rc;
        if ((rc = dev_alloc(ix, &ix->d_inline, n_slices * n_sites)) != ISSL_OK) return rc;
        if (residual_layout_fits(ix, &ix->res_shift) &&
            (rc = dev_alloc(ix, &ix->d_res, n_slices * n_sites)) != ISSL_OK)
            return rc;
    }
    if ((rc = dev_alloc(ix, &ix->d_lstart, lstart_count)) != ISSL_OK) return rc;
    if ((rc = dev_alloc(ix, &ix->d_meta, std::max<uint64_t>(n_slices, 1))) != ISSL_OK) return rc;
    if ((rc = dev_alloc(ix, &ix->d_mit64, kMitEntries)) != ISSL_OK) return rc;
    if ((rc = dev_alloc(ix, &ix->d_mit32, kMitEntries)) != ISSL_OK) return rc;
    // The tile kernels form MIT values from these factors: they must give the
    // table's values bit for bit (checked once on the host).
    if (!mit_factors_exact())
        return fail(ISSL_ERR_STATE, "MIT factor table does not reproduce the MIT closed form");
    if ((rc = dev_alloc(ix, &ix->d_mitf, kMitFactorEntries)) != ISSL_OK) return rc;
    if ((rc = dev_alloc(ix, &ix->d_cfd64, kCfdTable)) != ISSL_OK) return rc;
    if ((rc = dev_alloc(ix, &ix->d_cfd32, kCfdTable)) != ISSL_OK) return rc;
    ix->distinct_words = n_sites / 64 + 1;
    if ((rc = dev_alloc(ix, &ix->d_distinct, ix->distinct_words)) != ISSL_OK) return rc;
    if (n_slices)
        ISSL_HIP(hipMemcpy(ix->d_meta, ix->meta_h.data(), n_slices * sizeof(SliceMeta),
                           hipMemcpyHostToDevice));
    std::vector<float> m32(kMitEntries), c32(kCfdTable);
    std::vector<double> c64(kCfdTable, 1.0);  // [336] = 1.0: factor of an absent mismatch
    for (int r = 0; r < kMitEntries; ++r) m32[r] = static_cast<float>(mit_by_rank()[r]);
    std::memcpy(c64.data(), cfd_pos_table(), 320 * 8);
    std::memcpy(c64.data() + 320, cfd_pam_table(), 16 * 8);
    for (int i = 0; i < kCfdTable; ++i) c32[i] = static_cast<float>(c64[i]);
    ISSL_HIP(hipMemcpy(ix->d_mit64, mit_by_rank(), kMitEntries * 8, hipMemcpyHostToDevice));
    ISSL_HIP(hipMemcpy(ix->d_mit32, m32.data(), kMitEntries * 4, hipMemcpyHostToDevice));
    ISSL_HIP(hipMemcpy(ix->d_mitf, mit_factor_table(), kMitFactorEntries * 8, hipMemcpyHostToDevice));
    ISSL_HIP(hipMemcpy(ix->d_cfd64, c64.data(), kCfdTable * 8, hipMemcpyHostToDevice));
    ISSL_HIP(hipMemcpy(ix->d_cfd32, c32.data(), kCfdTable * 4, hipMemcpyHostToDevice));
    ISSL_HIP(hipMemset(ix->d_distinct, 0, ix->distinct_words * 8));
    return ISSL_OK;
}

// Per-slice list starts from the list sizes (ISSLScoringModule.cpp:181-198);
// sizes[s] holds slice s's 4^w list sizes.  Also sets p_max.
int build_lstart(issl_index *ix, const std::vector<const uint64_t *> &sizes,
                 std::vector<uint32_t> &lstart) {
    for (uint64_t s = 0; s < ix->S; ++s) {
        const SliceMeta &m = ix->meta_h[s];
        const uint64_t nk = 1ULL << (2 * m.width);
        const uint64_t *sz = sizes[s];
        uint64_t run = 0, longest = 0;
        for (uint64_t k = 0; k < nk; ++k) {
            lstart[m.lbase + k] = static_cast<uint32_t>(run);
            run += sz[k];
            longest = std::max<uint64_t>(longest, sz[k]);
        }
        if (run != ix->n)
            return fail(ISSL_ERR_FORMAT, "slice " + std::to_string(s) + " list sizes sum to " +
                                             std::to_string(run) + ", not N");
        lstart[m.lbase + nk] = static_cast<uint32_t>(run);
        ix->p_max += longest;
    }
    return ISSL_OK;
}

int check_index_args(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices, int device) {
    if (n_sites >= (1ULL << 32)) return fail(ISSL_ERR_FORMAT, "ids are 32-bit: N must be < 2^32");
    if (n_slices > static_cast<uint64_t>(kMaxSlices))
        return fail(ISSL_ERR_ARG, "at most 64 slices are supported");
    // The scorer encodes guides with 20 positions and takes the distance on
    // the whole XOR of guide and site signatures
    // (lib/ISSLScoringModule.cpp:219,286-290; ISSLScoreOfftargets.cpp:360-366),
    // so a site with more positions differs from every guide wherever those
    // extra codes are not A, and its MIT lookup fails (ISSLScoringModule.cpp:296).
    // The kernels take the distance on 20 positions; an index of any other
    // length is refused rather than scored differently.
    if (seq_len != 20)
        return fail(ISSL_ERR_FORMAT, "index sequence length is " + std::to_string(seq_len) +
                                         ": the scorer takes 20-position signatures only");
    const int ndev = issl_device_count();
    if (ndev <= 0) return fail(ISSL_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(ISSL_ERR_ARG, "device ordinal out of range");
    return ISSL_OK;
}

}  // namespace issl

extern "C" int issl_index_from_arrays(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                                      const uint64_t *sigs, const uint64_t *slice_masks,
                                      const uint64_t *list_sizes, const uint64_t *entries,
                                      int device, issl_index **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    *out = nullptr;
    if ((n_sites && !sigs) || (n_slices && (!slice_masks || !list_sizes)) ||
        (n_sites && n_slices && !entries))
        return fail(ISSL_ERR_ARG, "null index array");
    // The file's own consistency first, then the device.
    if (seq_len == 20 && n_sites < (1ULL << 32) && count_wide_sigs(sigs, n_sites) != 0)
        return fail(ISSL_ERR_FORMAT, "signatures hold codes beyond position 19");
    int rc = check_index_args(n_sites, seq_len, n_slices, device);
    if (rc != ISSL_OK) return rc;
    std::unique_ptr<issl_index, void (*)(issl_index *)> ix(new (std::nothrow) issl_index, free_index);
    if (!ix) return fail(ISSL_ERR_NOMEM, "host allocation failed");
    ix->device = device;
    uint64_t lstart_count = 0;
    if ((rc = init_meta(ix.get(), n_sites, seq_len, n_slices, slice_masks, &lstart_count)) != ISSL_OK)
        return rc;
    std::vector<uint32_t> lstart(lstart_count);
    std::vector<const uint64_t *> sizes(n_slices);
    for (uint64_t s = 0, base = 0; s < n_slices; ++s) {
        sizes[s] = list_sizes + base;
        base += 1ULL << (2 * ix->meta_h[s].width);
    }
    if ((rc = build_lstart(ix.get(), sizes, lstart)) != ISSL_OK) return rc;
    if (n_slices && count_bad_ids(entries, n_slices * n_sites, n_sites) != 0)
        return fail(ISSL_ERR_FORMAT, "posting entries hold ids >= N");

    DeviceGuard guard(device);
    if ((rc = alloc_index(ix.get(), lstart_count, true)) != ISSL_OK) return rc;
    if ((rc = upload(ix.get(), ix->d_sig, sigs, n_sites * 8)) != ISSL_OK) return rc;
    if ((rc = upload(ix.get(), ix->d_post, entries, n_slices * n_sites * 8)) != ISSL_OK) return rc;
    if ((rc = upload(ix.get(), ix->d_lstart, lstart.data(), lstart.size() * 4)) != ISSL_OK) return rc;
    if ((rc = build_inline_layout(ix.get())) != ISSL_OK) return rc;
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    // Release the staging buffer used for the upload.
    ix->ws.pinned.release();
    *out = ix.release();
    return ok();
}

extern "C" int issl_index_open(const char *path, int device, issl_index **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    uint64_t n, L, S, nl;
    int rc = issl_read_header(path, &n, &L, &S, &nl);
    if (rc != ISSL_OK) return rc;
    std::vector<uint64_t> sigs, masks, sizes, entries;
    try {
        sigs.resize(n);
        masks.resize(S);
        sizes.resize(nl);
        entries.resize(S * n);
    } catch (...) {
        return fail(ISSL_ERR_NOMEM, "host allocation for the index failed");
    }
    rc = issl_read_index(path, sigs.data(), masks.data(), sizes.data(), entries.data());
    if (rc != ISSL_OK) return rc;
    return issl_index_from_arrays(n, L, S, sigs.data(), masks.data(), sizes.data(), entries.data(),
                                  device, out);
}

// ------------------------------------------------------------ device build

namespace {

// Slice key and id of every site (ids are array positions).
__global__ void k_keys(const uint64_t *__restrict__ sig, uint64_t n, uint64_t fold,
                       uint32_t *__restrict__ keys, uint32_t *__restrict__ ids) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        keys[i] = slice_key(sig[i], fold);
        ids[i] = static_cast<uint32_t>(i);
    }
}

// List sizes of one slice.
__global__ void k_count(const uint32_t *__restrict__ keys, uint64_t n, uint32_t *__restrict__ count) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        atomicAdd(count + keys[i], 1u);
}

// Posting entries occ<<32 | id in list order.
__global__ void k_entries(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ occ,
                          uint64_t n, uint64_t *__restrict__ post) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t id = ids[i];
        post[i] = (static_cast<uint64_t>(occ[id]) << 32) | id;
    }
}

}  // namespace

namespace issl {

int index_build(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices, const uint64_t *sigs,
                const uint32_t *occ, bool on_device, const uint64_t *slice_masks, int device,
                issl_index **out) {
    if (!out) return fail(ISSL_ERR_ARG, "null output handle");
    *out = nullptr;
    if ((n_sites && (!sigs || !occ)) || (n_slices && !slice_masks))
        return fail(ISSL_ERR_ARG, "null index array");
    if (n_sites > 0x7FFFFFFFULL)
        return fail(ISSL_ERR_ARG, "device build takes at most 2^31 - 1 sites (build on the host)");
    if (!on_device && seq_len == 20 && count_wide_sigs(sigs, n_sites) != 0)
        return fail(ISSL_ERR_FORMAT, "signatures hold codes beyond position 19");
    int rc = check_index_args(n_sites, seq_len, n_slices, device);
    if (rc != ISSL_OK) return rc;
    std::unique_ptr<issl_index, void (*)(issl_index *)> ix(new (std::nothrow) issl_index, free_index);
    if (!ix) return fail(ISSL_ERR_NOMEM, "host allocation failed");
    ix->device = device;
    uint64_t lstart_count = 0;
    if ((rc = init_meta(ix.get(), n_sites, seq_len, n_slices, slice_masks, &lstart_count)) != ISSL_OK)
        return rc;
    DeviceGuard guard(device);
    if ((rc = alloc_index(ix.get(), lstart_count, true)) != ISSL_OK) return rc;
    hipStream_t st = ix->stream;
    if (on_device)
        ISSL_HIP(hipMemcpyAsync(ix->d_sig, sigs, n_sites * 8, hipMemcpyDeviceToDevice, st));
    else if ((rc = upload(ix.get(), ix->d_sig, sigs, n_sites * 8)) != ISSL_OK)
        return rc;
    const uint64_t n = n_sites;
    uint64_t max_nk = 1;
    int max_bits = 0;
    for (const SliceMeta &m : ix->meta_h) {
        max_nk = std::max<uint64_t>(max_nk, 1ULL << (2 * m.width));
        max_bits = std::max(max_bits, static_cast<int>(2 * m.width));
    }
    // Scratch: occurrences, keys and ids (and their sorted copies), list sizes,
    // the longest list, and the sort / scan / reduce temporaries.
    DevBuf b_occ, b_k, b_ks, b_i, b_is, b_cnt, b_tmp;
    if ((rc = b_occ.reserve(std::max<uint64_t>(n, 1) * 4)) != ISSL_OK) return rc;
    if (on_device)
        ISSL_HIP(hipMemcpyAsync(b_occ.p, occ, n * 4, hipMemcpyDeviceToDevice, st));
    else if ((rc = upload(ix.get(), b_occ.p, occ, n * 4)) != ISSL_OK)
        return rc;
    const size_t kb = std::max<uint64_t>(n, 1) * 4;
    if ((rc = b_k.reserve(kb)) != ISSL_OK || (rc = b_ks.reserve(kb)) != ISSL_OK ||
        (rc = b_i.reserve(kb)) != ISSL_OK || (rc = b_is.reserve(kb)) != ISSL_OK ||
        (rc = b_cnt.reserve((max_nk + 1) * 4 + 256)) != ISSL_OK)
        return rc;
    uint32_t *keys = static_cast<uint32_t *>(b_k.p), *keys_s = static_cast<uint32_t *>(b_ks.p);
    uint32_t *ids = static_cast<uint32_t *>(b_i.p), *ids_s = static_cast<uint32_t *>(b_is.p);
    uint32_t *cnt = static_cast<uint32_t *>(b_cnt.p);
    uint32_t *dmax = reinterpret_cast<uint32_t *>(static_cast<char *>(b_cnt.p) + (max_nk + 1) * 4);
    size_t t_sort = 0, t_scan = 0, t_max = 0;
    ISSL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, keys, keys_s, ids, ids_s,
                                                static_cast<int>(n), 0, std::max(max_bits, 1), st));
    ISSL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, cnt, ix->d_lstart,
                                              static_cast<int>(max_nk + 1), st));
    ISSL_HIP(hipcub::DeviceReduce::Max(nullptr, t_max, cnt, dmax, static_cast<int>(max_nk), st));
    const size_t t_bytes = std::max(t_sort, std::max(t_scan, t_max));
    if ((rc = b_tmp.reserve(t_bytes)) != ISSL_OK) return rc;
    const dim3 grid(4096), block(256);
    for (uint64_t s = 0; s < n_slices; ++s) {
        const SliceMeta &m = ix->meta_h[s];
        const uint64_t nk = 1ULL << (2 * m.width);
        size_t tb = t_bytes;
        hipLaunchKernelGGL(k_keys, grid, block, 0, st, ix->d_sig, n, m.fold, keys, ids);
        ISSL_HIP(hipMemsetAsync(cnt, 0, (nk + 1) * 4, st));
        hipLaunchKernelGGL(k_count, grid, block, 0, st, keys, n, cnt);
        // list starts with the end sentinel: the exclusive sum over nk + 1 counts, the last 0
        ISSL_HIP(hipcub::DeviceScan::ExclusiveSum(b_tmp.p, tb, cnt, ix->d_lstart + m.lbase,
                                                  static_cast<int>(nk + 1), st));
        tb = t_bytes;
        ISSL_HIP(hipcub::DeviceReduce::Max(b_tmp.p, tb, cnt, dmax, static_cast<int>(nk), st));
        uint32_t longest = 0;
        ISSL_HIP(hipMemcpyAsync(&longest, dmax, 4, hipMemcpyDeviceToHost, st));
        // LSD radix sort is stable: ids stay ascending inside each list
        // (ISSLCreateIndex.cpp:450-484 fills lists in id order).
        tb = t_bytes;
        ISSL_HIP(hipcub::DeviceRadixSort::SortPairs(b_tmp.p, tb, keys, keys_s, ids, ids_s,
                                                    static_cast<int>(n), 0,
                                                    std::max(static_cast<int>(2 * m.width), 1), st));
        hipLaunchKernelGGL(k_entries, grid, block, 0, st, ids_s, static_cast<const uint32_t *>(b_occ.p),
                           n, ix->d_post + s * n);
        ISSL_HIP(hipGetLastError());
        ISSL_HIP(hipStreamSynchronize(st));
        ix->p_max += longest;
    }
    if ((rc = build_inline_layout(ix.get())) != ISSL_OK) return rc;
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    ix->ws.pinned.release();
    *out = ix.release();
    return ISSL_OK;
}

}  // namespace issl

extern "C" int issl_index_build(uint64_t n_sites, uint64_t seq_len, uint64_t n_slices,
                                const uint64_t *sigs, const uint32_t *occ,
                                const uint64_t *slice_masks, int device, issl_index **out) {
    const int rc = index_build(n_sites, seq_len, n_slices, sigs, occ, false, slice_masks, device, out);
    return rc == ISSL_OK ? ok() : rc;
}

extern "C" int issl_index_export(issl_index *ix, uint64_t *list_sizes, uint64_t *entries) {
    if (!ix || (ix->n_lists && !list_sizes)) return fail(ISSL_ERR_ARG, "null argument");
    if (ix->src) return fail(ISSL_ERR_ARG, "export needs a resident index");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    uint64_t lstart_count = 0;
    for (const SliceMeta &m : ix->meta_h) lstart_count += (1ULL << (2 * m.width)) + 1;
    std::vector<uint32_t> lstart(lstart_count);
    ISSL_HIP(hipMemcpy(lstart.data(), ix->d_lstart, lstart_count * 4, hipMemcpyDeviceToHost));
    uint64_t k = 0;
    for (const SliceMeta &m : ix->meta_h) {
        const uint64_t nk = 1ULL << (2 * m.width);
        for (uint64_t j = 0; j < nk; ++j) list_sizes[k++] = lstart[m.lbase + j + 1] - lstart[m.lbase + j];
    }
    if (entries && ix->S * ix->n)
        ISSL_HIP(hipMemcpy(entries, ix->d_post, ix->S * ix->n * 8, hipMemcpyDeviceToHost));
    return ok();
}

extern "C" int issl_index_close(issl_index *ix) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    free_index(ix);
    return ok();
}

extern "C" int issl_index_get_info(const issl_index *ix, issl_index_info *out) {
    if (!ix || !out) return fail(ISSL_ERR_ARG, "null argument");
    out->n_sites = ix->n;
    out->seq_len = ix->seq_len;
    out->n_slices = ix->S;
    out->n_lists = ix->n_lists;
    out->device_bytes = ix->device_bytes;
    out->device = ix->device;
    out->max_width = ix->max_w;
    return ok();
}

extern "C" int issl_index_set_residual(issl_index *ix, int32_t use) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    ix->res_use = use != 0;
    return ok();
}

extern "C" int issl_index_get_residual(const issl_index *ix, int32_t *built, int32_t *in_use,
                                       uint32_t *occ_bits) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    const bool b = ix->d_res != nullptr;
    if (built) *built = b ? 1 : 0;
    if (in_use) *in_use = (b && ix->res_use) ? 1 : 0;
    if (occ_bits) *occ_bits = b ? 32u - ix->res_shift : 0u;
    return ok();
}

extern "C" int issl_index_validate(issl_index *ix, uint64_t *bad_entries) {
    if (!ix || !bad_entries) return fail(ISSL_ERR_ARG, "null argument");
    if (ix->src) return fail(ISSL_ERR_ARG, "validation needs a resident index (a streamed index "
                                           "checks its posting ids as it copies them)");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    int rc = ix->ws.scratch.reserve(ix->n * 4 + 64);
    if (rc != ISSL_OK) return rc;
    uint32_t *seen = static_cast<uint32_t *>(ix->ws.scratch.p);
    unsigned long long *bad = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(ix->ws.scratch.p) + ((ix->n * 4 + 15) / 16) * 16);
    ISSL_HIP(hipMemsetAsync(bad, 0, 8, ix->stream));
    for (uint64_t s = 0; s < ix->S; ++s) {
        ISSL_HIP(hipMemsetAsync(seen, 0, ix->n * 4, ix->stream));
        hipLaunchKernelGGL(k_validate, dim3(2048), dim3(256), 0, ix->stream, ix->d_post, ix->d_sig,
                           ix->d_lstart, ix->meta_h[s], s, ix->n, seen, bad);
        ISSL_HIP(hipGetLastError());
    }
    unsigned long long h = 0;
    ISSL_HIP(hipMemcpyAsync(&h, bad, 8, hipMemcpyDeviceToHost, ix->stream));
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    *bad_entries = h;
    return ok();
}

extern "C" int issl_distinct_count(issl_index *ix, uint64_t *out) {
    if (!ix || !out) return fail(ISSL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    std::vector<uint64_t> words(ix->distinct_words);
    ISSL_HIP(hipMemcpyAsync(words.data(), ix->d_distinct, ix->distinct_words * 8,
                            hipMemcpyDeviceToHost, ix->stream));
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    uint64_t c = 0;
    for (uint64_t w : words) c += static_cast<uint64_t>(__builtin_popcountll(w));
    *out = c;
    return ok();
}

extern "C" int issl_distinct_reset(issl_index *ix) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    ISSL_HIP(hipMemsetAsync(ix->d_distinct, 0, ix->distinct_words * 8, ix->stream));
    ISSL_HIP(hipStreamSynchronize(ix->stream));
    return ok();
}

namespace issl {

int ws_acquire(issl_index *ix, hipStream_t st) {
    if (!ix->ws_used || ix->ws_stream == st) return ISSL_OK;
    if (ix->ws_lazy) {
        // Everything enqueued on the index's stream so far, the last use of
        // the workspace included.
        if (!ix->ws_done) ISSL_HIP(hipEventCreateWithFlags(&ix->ws_done, hipEventDisableTiming));
        ISSL_HIP(hipEventRecord(ix->ws_done, ix->ws_stream));
        ix->ws_lazy = false;
    }
    ISSL_HIP(hipStreamWaitEvent(st, ix->ws_done, 0));
    return ISSL_OK;
}

int ws_release(issl_index *ix, hipStream_t st) {
    ix->ws_used = true;
    ix->ws_stream = st;
    ix->ws_lazy = st == ix->stream;
    if (ix->ws_lazy) return ISSL_OK;
    if (!ix->ws_done) ISSL_HIP(hipEventCreateWithFlags(&ix->ws_done, hipEventDisableTiming));
    ISSL_HIP(hipEventRecord(ix->ws_done, st));
    return ISSL_OK;
}

}  // namespace issl

namespace {

int score_on_device(issl_index *ix, const uint64_t *d_guides, uint64_t n, const issl_params *p,
                    const ScoreScalars &sc, const DevResults &o, hipStream_t st,
                    Workspace *ws = nullptr) {
    uint64_t *distinct = (p->flags & ISSL_FLAG_DISTINCT) ? ix->d_distinct : nullptr;
    if (ix->src) return score_streamed(ix, d_guides, n, p, sc, o, distinct, st);
    return score_with(ix, d_guides, n, p, sc, ix->tuning, o, ws ? *ws : ix->ws, st);
}

// The workspace of a bound stream, or null.
Workspace *bound_ws(issl_index *ix, hipStream_t st) {
    for (auto &b : ix->bound)
        if (b->stream == st) return &b->ws;
    return nullptr;
}

// Whether a captured sequence's workspace is still allocated (the index's
// own workspace or a bound stream's, at its current address).
bool ws_live(const issl_index *ix, const void *tiles) {
    if (tiles == ix->ws.tiles.p) return true;
    for (const auto &b : ix->bound)
        if (tiles == b->ws.tiles.p) return true;
    return false;
}

}  // namespace

namespace issl {

int params_to_scalars(const issl_index *ix, const issl_params *p, ScoreScalars *sc) {
    return check_params(ix, p, sc);
}

int score_with(issl_index *ix, const uint64_t *d_guides, uint64_t n, const issl_params *p,
               const ScoreScalars &sc, const issl_tuning &tn, const DevResults &o, Workspace &ws,
               hipStream_t st) {
    uint64_t *distinct = (p->flags & ISSL_FLAG_DISTINCT) ? ix->d_distinct : nullptr;
    if (ix->src) return fail(ISSL_ERR_ARG, "a streamed index is scored through issl_score or issl_score_device");
    // Launches on a bound stream run beside other streams' launches: the tile
    // kernels take ovl_grid workgroups per CU (carried to the launchers in the
    // tuning's ovl_grid; other launches, the autotuner's included, carry 0).
    bool on_bound = false;
    for (const auto &b : ix->bound) on_bound = on_bound || b->stream == st;
    issl_tuning t = tn;
    t.ovl_grid = on_bound ? (ix->ovl_grid ? ix->ovl_grid : kOvlGridDefault) : 0;
    if (p->form == ISSL_FORM_ORDERED && ordered_fits(ix))
        return launch_ordered(ix, d_guides, n, sc, t, o, distinct, ws, st);
    if (p->form != ISSL_FORM_BATCHED)
        return launch_plain(ix, d_guides, n, sc, o, distinct, nullptr, nullptr, st);
    return launch_batched(ix, d_guides, n, sc, p->precision, p->chunk, t, o, distinct, ws, st);
}

}  // namespace issl

extern "C" int issl_score(issl_index *ix, const uint64_t *guides, uint64_t n,
                          const issl_params *params, issl_results *out) {
    ScoreScalars sc;
    int rc = check_params(ix, params, &sc);
    if (rc != ISSL_OK) return rc;
    if (!out) return fail(ISSL_ERR_ARG, "null results");
    if (n == 0) return ok();
    if (!guides) return fail(ISSL_ERR_ARG, "null guides");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    // A resident index scores the guides in chunks of kHostChunk, each with
    // its guides copied into pinned staging, sent, scored and its results sent
    // back on the index's stream: the host stages chunk j + 1 and copies out
    // chunk j's results while the device works on later chunks.  A streamed
    // index passes all its slices through HBM once per launch, so it scores
    // the call in one piece.  Only the columns the caller asked for are
    // written and copied back (the pipeline module asks for MIT, CFD and the
    // call: 17 of the 49 bytes per guide); each chunk's columns lie packed
    // together at j * L.total in the output buffers.
    const uint64_t C = ix->src ? n : std::min<uint64_t>(n, kHostChunk);
    const uint64_t nch = (n + C - 1) / C;
    const OutLayout L(C, *out);
    if ((rc = ix->ws.guides.reserve(n * 8)) != ISSL_OK) return rc;
    if ((rc = ix->ws.out.reserve(nch * L.total)) != ISSL_OK) return rc;
    if ((rc = ix->ws.pinned.reserve(n * 8 + nch * L.total)) != ISSL_OK) return rc;
    char *pin = static_cast<char *>(ix->ws.pinned.p);
    char *hout = pin + n * 8;
    char *ob = static_cast<char *>(ix->ws.out.p);
    uint64_t *dg = static_cast<uint64_t *>(ix->ws.guides.p);
    hipStream_t st = ix->stream;
    // chunk j's results are on the host once done[j] has passed
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } done;
    if (nch > 1) {
        done.e.assign(nch, nullptr);
        for (hipEvent_t &x : done.e) ISSL_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    }
    if ((rc = ws_acquire(ix, st)) != ISSL_OK) return rc;
    for (uint64_t j = 0; j < nch; ++j) {
        const uint64_t g0 = j * C, cn = std::min(C, n - g0);
        std::memcpy(pin + g0 * 8, guides + g0, cn * 8);
        ISSL_HIP(hipMemcpyAsync(dg + g0, pin + g0 * 8, cn * 8, hipMemcpyHostToDevice, st));
        const OutLayout Lj(cn, *out);
        if ((rc = score_on_device(ix, dg + g0, cn, params, sc, Lj.bind(ob + j * L.total), st)) != ISSL_OK)
            return rc;
        ISSL_HIP(hipMemcpyAsync(hout + j * L.total, ob + j * L.total, Lj.total, hipMemcpyDeviceToHost, st));
        if (nch > 1) ISSL_HIP(hipEventRecord(done.e[j], st));
    }
    if ((rc = ws_release(ix, st)) != ISSL_OK) return rc;
    if (nch == 1) ISSL_HIP(hipStreamSynchronize(st));
    for (uint64_t j = 0; j < nch; ++j) {
        if (nch > 1) ISSL_HIP(hipEventSynchronize(done.e[j]));
        const uint64_t g0 = j * C, cn = std::min(C, n - g0);
        const OutLayout Lj(cn, *out);
        const char *h = hout + j * L.total;
        if (out->mit) std::memcpy(out->mit + g0, h + Lj.mit, cn * 8);
        if (out->cfd) std::memcpy(out->cfd + g0, h + Lj.cfd, cn * 8);
        if (out->mit_sum) std::memcpy(out->mit_sum + g0, h + Lj.msum, cn * 8);
        if (out->cfd_sum) std::memcpy(out->cfd_sum + g0, h + Lj.csum, cn * 8);
        if (out->occ_count) std::memcpy(out->occ_count + g0, h + Lj.occ, cn * 8);
        if (out->site_count) std::memcpy(out->site_count + g0, h + Lj.sites, cn * 8);
        if (out->pass) std::memcpy(out->pass + g0, h + Lj.pass, cn);
    }
    return ok();
}

extern "C" int issl_score_device(issl_index *ix, const uint64_t *guides_device, uint64_t n,
                                 const issl_params *params, issl_results *out_device, void *stream) {
    ScoreScalars sc;
    int rc = check_params(ix, params, &sc);
    if (rc != ISSL_OK) return rc;
    if (!out_device) return fail(ISSL_ERR_ARG, "null results");
    if (n == 0) return ok();
    if (!guides_device) return fail(ISSL_ERR_ARG, "null guides");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    DevResults o;
    o.mit = out_device->mit;
    o.cfd = out_device->cfd;
    o.mit_sum = out_device->mit_sum;
    o.cfd_sum = out_device->cfd_sum;
    o.occ = out_device->occ_count;
    o.sites = out_device->site_count;
    o.pass = out_device->pass;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
    if ((params->flags & ISSL_FLAG_GRAPH) && ix->src)
        return fail(ISSL_ERR_ARG, "graph replay needs a resident index (a streamed index copies "
                                  "postings from the host during the call)");
    // A bound stream scores in its own workspace; any other stream shares the
    // index's, after the work of the stream that used it last.
    Workspace *const bws = bound_ws(ix, st);
    Workspace &W = bws ? *bws : ix->ws;
    if (!bws && (rc = ws_acquire(ix, st)) != ISSL_OK) return rc;
    auto release = [&]() -> int { return bws ? ISSL_OK : ws_release(ix, st); };
    const bool use_graph = (params->flags & ISSL_FLAG_GRAPH) && !ix->timing;
    if (use_graph) {
        // Replay a captured sequence made for exactly this call.  Entries made
        // for a workspace address no longer allocated hold freed pointers:
        // drop them.
        const void *wsp = W.tiles.p;
        for (size_t i = 0; i < ix->graphs.size();) {
            GraphEntry &gr = ix->graphs[i];
            if (!ws_live(ix, gr.ws)) {
                (void)hipGraphExecDestroy(gr.exec);
                ix->graphs.erase(ix->graphs.begin() + static_cast<long>(i));
                continue;
            }
            if (gr.ws == wsp && gr.guides == guides_device && gr.n == n && gr.stream == st &&
                std::memcmp(&gr.params, params, sizeof(issl_params)) == 0 &&
                std::memcmp(&gr.out, out_device, sizeof(issl_results)) == 0 &&
                std::memcmp(&gr.tuning, &ix->tuning, sizeof(issl_tuning)) == 0) {
                ISSL_HIP(hipGraphLaunch(gr.exec, st));
                if ((rc = release()) != ISSL_OK) return rc;
                return ok();
            }
            ++i;
        }
    }
    rc = score_on_device(ix, guides_device, n, params, sc, o, st, &W);
    if (rc != ISSL_OK) return rc;
    if (use_graph) {
        // The workspace is in place now: capture the same sequence for replay.
        hipGraph_t graph = nullptr;
        ISSL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = score_on_device(ix, guides_device, n, params, sc, o, st, &W);
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != ISSL_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc != ISSL_OK ? rc : fail(ISSL_ERR_DEVICE, std::string("graph capture: ") +
                                                                  hipGetErrorString(e));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess)
            return fail(ISSL_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
        if (ix->graphs.size() >= 8) {  // keep the eight most recent
            (void)hipGraphExecDestroy(ix->graphs.front().exec);
            ix->graphs.erase(ix->graphs.begin());
        }
        ix->graphs.push_back(GraphEntry{guides_device, n, *params, *out_device, st, ix->tuning,
                                        W.tiles.p, exec});
    }
    if ((rc = release()) != ISSL_OK) return rc;
    return ok();
}

extern "C" int issl_index_bind_stream(issl_index *ix, void *stream) {
    if (!ix || !stream) return fail(ISSL_ERR_ARG, "null argument");
    if (ix->src)
        return fail(ISSL_ERR_ARG, "a streamed index stages its postings through one set of "
                                  "buffers: its calls cannot overlap");
    std::lock_guard<std::mutex> lk(ix->mu);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (st == ix->stream) return fail(ISSL_ERR_ARG, "the index's own stream uses the index's workspace");
    if (bound_ws(ix, st)) return ok();
    if (ix->bound.size() >= static_cast<size_t>(kMaxBoundStreams))
        return fail(ISSL_ERR_CAPACITY, "at most " + std::to_string(kMaxBoundStreams) +
                                           " streams bound to one index");
    auto b = std::make_unique<issl_index::BoundWs>();
    b->stream = st;
    ix->bound.push_back(std::move(b));
    return ok();
}

extern "C" int issl_index_unbind_stream(issl_index *ix, void *stream) {
    if (!ix || !stream) return fail(ISSL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (size_t i = 0; i < ix->bound.size(); ++i) {
        if (ix->bound[i]->stream != st) continue;
        // The stream's queued calls use the workspace until they finish.
        ISSL_HIP(hipStreamSynchronize(st));
        const void *wsp = ix->bound[i]->ws.tiles.p;
        for (size_t j = 0; j < ix->graphs.size();) {
            if (ix->graphs[j].ws == wsp && wsp) {
                (void)hipGraphExecDestroy(ix->graphs[j].exec);
                ix->graphs.erase(ix->graphs.begin() + static_cast<long>(j));
                continue;
            }
            ++j;
        }
        ix->bound.erase(ix->bound.begin() + static_cast<long>(i));
        return ok();
    }
    return fail(ISSL_ERR_ARG, "stream not bound to this index");
}

extern "C" int issl_candidates(issl_index *ix, const uint64_t *guides, uint64_t n, uint64_t *offsets,
                               uint32_t *ids, uint64_t capacity) {
    if (!ix || !offsets) return fail(ISSL_ERR_ARG, "null argument");
    if (n && !guides) return fail(ISSL_ERR_ARG, "null guides");
    offsets[0] = 0;
    if (n == 0) return ok();
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    int rc;
    if ((rc = ix->ws.guides.reserve(n * 8)) != ISSL_OK) return rc;
    if ((rc = ix->ws.out.reserve(n * 8 * 2 + 64)) != ISSL_OK) return rc;
    hipStream_t st = ix->stream;
    if ((rc = ws_acquire(ix, st)) != ISSL_OK) return rc;
    ISSL_HIP(hipMemcpyAsync(ix->ws.guides.p, guides, n * 8, hipMemcpyHostToDevice, st));
    ScoreScalars sc{};
    sc.method = ISSL_METHOD_MIT_AND_CFD;
    sc.calc_mit = 0;
    sc.calc_cfd = 0;
    sc.use_cut = 0;
    sc.threshold = 0;
    DevResults o;
    o.sites = static_cast<uint64_t *>(ix->ws.out.p);
    const uint64_t *dg = static_cast<const uint64_t *>(ix->ws.guides.p);
    if ((rc = launch_plain(ix, dg, n, sc, o, nullptr, nullptr, nullptr, st)) != ISSL_OK) return rc;
    std::vector<uint64_t> counts(n);
    ISSL_HIP(hipMemcpyAsync(counts.data(), o.sites, n * 8, hipMemcpyDeviceToHost, st));
    ISSL_HIP(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + counts[i];
    const uint64_t total = offsets[n];
    if (total > capacity)
        return fail(ISSL_ERR_CAPACITY, "candidate capacity " + std::to_string(capacity) +
                                           " < " + std::to_string(total));
    if (total == 0) return ok();
    if (!ids) return fail(ISSL_ERR_ARG, "null ids");
    if ((rc = ix->ws.scratch.reserve(total * 4 + (n + 1) * 8)) != ISSL_OK) return rc;
    uint64_t *d_off = static_cast<uint64_t *>(ix->ws.scratch.p);
    uint32_t *d_ids = reinterpret_cast<uint32_t *>(d_off + n + 1);
    ISSL_HIP(hipMemcpyAsync(d_off, offsets, (n + 1) * 8, hipMemcpyHostToDevice, st));
    if ((rc = launch_plain(ix, dg, n, sc, o, nullptr, d_off, d_ids, st)) != ISSL_OK) return rc;
    ISSL_HIP(hipMemcpyAsync(ids, d_ids, total * 4, hipMemcpyDeviceToHost, st));
    ISSL_HIP(hipStreamSynchronize(st));
    return ok();
}

extern "C" int issl_timing_enable(issl_index *ix, int enable) {
    if (!ix) return fail(ISSL_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    for (size_t i = 0; i < ix->ev_used; ++i) ISSL_HIP(hipEventSynchronize(ix->ev_pool[i]));
    ix->ev_used = 0;
    if (enable < 0 || enable > 1024) return fail(ISSL_ERR_ARG, "timing repeat must be 0..1024");
    ix->timing = enable;
    return ok();
}

extern "C" int issl_timing_read(issl_index *ix, double *kernel_ms, uint64_t *launches) {
    if (!ix || !kernel_ms || !launches) return fail(ISSL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(ix->mu);
    DeviceGuard guard(ix->device);
    double total = 0;
    for (size_t i = 0; i + 1 < ix->ev_used; i += 2) {
        ISSL_HIP(hipEventSynchronize(ix->ev_pool[i + 1]));
        float ms = 0;
        ISSL_HIP(hipEventElapsedTime(&ms, ix->ev_pool[i], ix->ev_pool[i + 1]));
        total += ms;
    }
    *kernel_ms = total;
    *launches = ix->ev_used / 2 * static_cast<uint64_t>(std::max(ix->timing, 1));
    return ok();
}
