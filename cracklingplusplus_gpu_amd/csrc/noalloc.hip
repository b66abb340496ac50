// noalloc.hip -- the allocation-free scoring entry point.
//
// issl_score_device_noalloc is issl_score_device with the BATCHED / ORDERED
// scoring workspace (tile plans, per-tile partials, the tile queue and its
// counters) carved from `scratch`, a file-scope device array of kScratchBytes,
// instead of device memory the library allocates.  It calls no hipMalloc,
// hipFree or allocating routine, so callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid it) can score.
//
// Each device holds its own copy of `scratch` at its own address, kept until
// the process ends; the host obtains that address once per device with
// hipGetSymbolAddress (the array's name in host code is only a stand-in).
// Every call on a device shares its copy: the tile queue's counters are
// cleared by a fill kernel at the start of each call, and calls that overlap
// on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "device.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up on the first call there.
int scratch_on(int device, char **out) {
    if (device < 0 || device >= kMaxDevices) return issl::fail(ISSL_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        issl::DeviceGuard guard(device);
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return issl::fail(ISSL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                                   hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[device];
    return ISSL_OK;
}

// Most guides whose workspace fits the scratch array for this index.
uint64_t max_guides(const issl_index *ix, bool res) {
    uint64_t lo = 0, hi = 1ULL << 20;
    const int S = static_cast<int>(ix->S);
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (issl::batch_ws_bytes(mid, S, res) <= kScratchBytes)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace

extern "C" int issl_score_device_noalloc(issl_index *ix, const uint64_t *guides_device, uint64_t n,
                                         const issl_params *params, issl_results *out_device,
                                         void *stream) {
    issl::ScoreScalars sc;
    int rc = issl::params_to_scalars(ix, params, &sc);
    if (rc != ISSL_OK) return rc;
    if (!out_device) return issl::fail(ISSL_ERR_ARG, "null results");
    if (ix->src) return issl::fail(ISSL_ERR_ARG, "the allocation-free entry point needs a resident index");
    if (params->flags & ISSL_FLAG_GRAPH)
        return issl::fail(ISSL_ERR_ARG, "the allocation-free entry point does not capture graphs");
    if (n == 0) return issl::ok();
    if (!guides_device) return issl::fail(ISSL_ERR_ARG, "null guides");
    std::lock_guard<std::mutex> lk(ix->mu);
    // PLAIN (and ORDERED where it runs the PLAIN kernel) uses no workspace.
    const bool tiles = params->form == ISSL_FORM_BATCHED ||
                       (params->form == ISSL_FORM_ORDERED && issl::ordered_fits(ix));
    const bool res = ix->res_use && ix->d_res;
    if (tiles && issl::batch_ws_bytes(n, static_cast<int>(ix->S), res) > kScratchBytes)
        return issl::fail(ISSL_ERR_CAPACITY,
                          std::to_string(n) + " guides exceed the allocation-free workspace (" +
                              std::to_string(kScratchBytes >> 20) + " MiB per device): at most " +
                              std::to_string(max_guides(ix, res)) + " per call on this index");
    issl::DeviceGuard guard(ix->device);
    char *base = nullptr;
    if ((rc = scratch_on(ix->device, &base)) != ISSL_OK) return rc;
    issl::Workspace ws;
    ws.ext = base;
    ws.ext_bytes = kScratchBytes;
    issl::DevResults o;
    o.mit = out_device->mit;
    o.cfd = out_device->cfd;
    o.mit_sum = out_device->mit_sum;
    o.cfd_sum = out_device->cfd_sum;
    o.occ = out_device->occ_count;
    o.sites = out_device->site_count;
    o.pass = out_device->pass;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
    rc = issl::score_with(ix, guides_device, n, params, sc, ix->tuning, o, ws, st);
    return rc == ISSL_OK ? issl::ok() : rc;
}
