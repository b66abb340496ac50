// server.hip -- request loop over resident indexes: admission by free HBM,
// continuous batching into in-flight slots, per-request latency.
//
// A server owns its indexes and `slots` launch slots.  A slot is a stream
// with its own device buffers, scoring workspace and pinned staging, so
// several merged batches are in flight at once: while one slot's kernels run,
// the next slot's guides are already copied in and its kernels queued behind
// or beside them.  Two threads drive it:
//   launcher   : takes the oldest request with guides not yet launched, fills
//                a free slot with up to max_batch guides from it and from every
//                queued request for the same index and parameters (requests
//                larger than a slot are split over several launches), and
//                queues copy-in, scoring and copy-out on the slot's stream;
//   completion : retires slots in launch order, scatters each slot's results
//                into the requests' host buffers and completes a request when
//                all its guides are back.
// A request is complete only when none of its pieces is still in a slot, also
// when it fails (a failed piece, index removal, server shutdown): the slots
// still holding its other pieces write into its buffers when they retire, so
// the caller may release those buffers only after the request completes.
// Each guide's result depends only on the guide, so how requests are merged
// or split never changes a result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device.hpp"
#include "issl_index.h"
#include "issl_server.h"

using namespace issl;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int32_t index_id = 0;
    std::vector<uint64_t> guides;
    issl_params params{};
    ScoreScalars sc{};
    issl_results out{};
    Clock::time_point t_submit;
    uint64_t launched = 0;      // guides handed to slots so far
    uint64_t returned = 0;      // guides whose results are back
    uint64_t pending = 0;       // pieces in slots not yet retired
    bool done = false;
    int32_t status = ISSL_OK;
    std::string error;
};

// A run of one request's guides inside a slot's batch.
struct Piece {
    std::shared_ptr<Request> req;
    uint64_t req_off, n;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf guides, out;
    Workspace ws;
    HostBuf pin;                 // guides in, then result columns out
    std::vector<Piece> pieces;
    uint64_t n = 0;
    uint64_t cap_n = 0;          // largest batch this slot has held
    int32_t index_id = -1;
    int rc = ISSL_OK;
    std::string error;
};

// Slot status while the launcher is still queueing its work (status codes are <= 0).
constexpr int kFilling = 1;

// Device bytes a slot holds per guide of its batch: guides, result columns,
// and the scoring workspace (tile partials, plans).
constexpr uint64_t kSlotBytesPerGuide = 8 + 57 + 64 * 32 + 65 * 4 + 64 * 8 + 8;

bool same_params(const issl_params &a, const issl_params &b) {
    return a.max_dist == b.max_dist && a.method == b.method && a.threshold == b.threshold &&
           a.form == b.form && a.precision == b.precision && a.flags == b.flags &&
           a.chunk == b.chunk;
}

template <typename T>
void put(T *dst, const char *col, uint64_t off, uint64_t n) {
    if (dst) std::memcpy(dst, col + off * sizeof(T), n * sizeof(T));
}

}  // namespace

struct issl_server {
    issl_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_launch;   // launcher: work and a free slot
    std::condition_variable cv_retire;   // completion: a slot in flight
    std::condition_variable cv_done;     // waiters and index removal
    std::deque<std::shared_ptr<Request>> queue;   // requests with guides not yet launched
    std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
    std::map<int32_t, issl_index *> indexes;
    std::map<int32_t, uint64_t> index_bytes;
    std::map<int32_t, int> index_inflight;        // slots in flight per index
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<Slot *> free_slots, inflight;
    uint64_t next_ticket = 1;
    int32_t next_index = 0;
    bool stop = false;
    std::thread launcher, completer;
    // statistics since the last reset
    std::vector<double> lat_ms;
    uint64_t done = 0, failed = 0, guides = 0, launches = 0, peak_inflight = 0;

    void launch_loop();
    void complete_loop();
    void fail_request(const std::shared_ptr<Request> &r, int rc, const std::string &msg);
    void settle(const std::shared_ptr<Request> &r);
    int fill(Slot *s, issl_index *ix, const issl_params &p, const ScoreScalars &sc,
             const issl_tuning &tn);
};

// Record a request's failure (the first one wins) and complete it if no slot
// still holds a piece of it (the caller holds mu and has taken it off the
// queue).
void issl_server::fail_request(const std::shared_ptr<Request> &r, int rc, const std::string &msg) {
    if (r->done) return;
    if (r->status == ISSL_OK) {
        r->status = rc;
        r->error = msg;
    }
    settle(r);
}

// Complete a request once no slot holds a piece of it and it has failed or
// all its guides are back (the caller holds mu).  Waiters are woken by the
// caller.
void issl_server::settle(const std::shared_ptr<Request> &r) {
    if (r->done || r->pending) return;
    if (r->status == ISSL_OK && r->returned < r->guides.size()) return;
    r->done = true;
    ++done;
    if (r->status != ISSL_OK) ++failed;
    guides += r->guides.size();
    lat_ms.push_back(
        std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count());
}

// Stage the slot's guides and queue its scoring on the slot's stream.
int issl_server::fill(Slot *s, issl_index *ix, const issl_params &p, const ScoreScalars &sc,
                      const issl_tuning &tn) {
    const uint64_t n = s->n;
    const OutLayout L(n);
    // Admission by free memory: growing a slot needs that much HBM beyond the reserve.
    const uint64_t grow = n > s->cap_n ? (n - s->cap_n) * kSlotBytesPerGuide : 0;
    if (grow) {
        size_t fr = 0, tot = 0;
        ISSL_HIP(hipMemGetInfo(&fr, &tot));
        if (grow + cfg.hbm_reserve > fr)
            return fail(ISSL_ERR_NOMEM, "admission refused: a batch of " + std::to_string(n) +
                                            " guides needs " + std::to_string(grow) +
                                            " B of HBM beyond the reserve");
    }
    int rc;
    if ((rc = s->guides.reserve(n * 8)) != ISSL_OK) return rc;
    if ((rc = s->out.reserve(L.total)) != ISSL_OK) return rc;
    if ((rc = s->pin.reserve(n * 8 + L.total)) != ISSL_OK) return rc;
    s->cap_n = std::max(s->cap_n, n);
    char *pin = static_cast<char *>(s->pin.p);
    uint64_t off = 0;
    for (const Piece &pc : s->pieces) {
        std::memcpy(pin + off * 8, pc.req->guides.data() + pc.req_off, pc.n * 8);
        off += pc.n;
    }
    ISSL_HIP(hipMemcpyAsync(s->guides.p, pin, n * 8, hipMemcpyHostToDevice, s->stream));
    const DevResults o = L.bind(static_cast<char *>(s->out.p));
    if ((rc = score_with(ix, static_cast<const uint64_t *>(s->guides.p), n, &p, sc, tn, o, s->ws,
                         s->stream)) != ISSL_OK)
        return rc;
    ISSL_HIP(hipMemcpyAsync(pin + n * 8, s->out.p, L.total, hipMemcpyDeviceToHost, s->stream));
    ISSL_HIP(hipEventRecord(s->done, s->stream));
    return ISSL_OK;
}

void issl_server::launch_loop() {
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_launch.wait(lk, [&] { return stop || (!queue.empty() && !free_slots.empty()); });
        if (stop) break;
        Slot *s = free_slots.front();
        free_slots.pop_front();
        // The oldest request with guides left, then every queued request it can share with.
        const std::shared_ptr<Request> head = queue.front();
        s->pieces.clear();
        s->n = 0;
        for (auto it = queue.begin(); it != queue.end() && s->n < cfg.max_batch;) {
            const std::shared_ptr<Request> r = *it;
            if (r->index_id != head->index_id || !same_params(r->params, head->params)) {
                ++it;
                continue;
            }
            const uint64_t take = std::min<uint64_t>(r->guides.size() - r->launched, cfg.max_batch - s->n);
            s->pieces.push_back(Piece{r, r->launched, take});
            r->launched += take;
            ++r->pending;
            s->n += take;
            if (r->launched == r->guides.size())
                it = queue.erase(it);
            else
                ++it;
        }
        s->index_id = head->index_id;
        s->rc = kFilling;
        s->error.clear();
        issl_index *ix = nullptr;
        auto ix_it = indexes.find(head->index_id);
        if (ix_it != indexes.end()) ix = ix_it->second;
        ++index_inflight[s->index_id];
        ++launches;
        inflight.push_back(s);
        peak_inflight = std::max<uint64_t>(peak_inflight, inflight.size());
        const issl_params p = head->params;
        const ScoreScalars sc = head->sc;
        lk.unlock();

        int rc = ISSL_ERR_STATE;
        std::string msg = "index was removed";
        if (ix) {
            issl_tuning tn;
            {
                std::lock_guard<std::mutex> ilk(ix->mu);
                tn = ix->tuning;
            }
            rc = fill(s, ix, p, sc, tn);
            msg = rc == ISSL_OK ? std::string() : std::string(last_error_cstr());
        }

        lk.lock();
        s->rc = rc;
        s->error = msg;
        cv_retire.notify_one();
    }
}

void issl_server::complete_loop() {
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_retire.wait(lk, [&] { return !inflight.empty() || stop; });
        if (inflight.empty()) break;  // stop requested and nothing in flight
        Slot *s = inflight.front();
        // The launcher may still be filling this slot (copy-in queued, kernels
        // being launched); its status is final when it leaves kFilling.
        cv_retire.wait(lk, [&] { return s->rc != kFilling; });
        int rc = s->rc;
        lk.unlock();
        if (rc == ISSL_OK) {
            const hipError_t e = hipEventSynchronize(s->done);
            if (e != hipSuccess) {
                rc = ISSL_ERR_DEVICE;
                s->error = std::string("scoring launch failed: ") + hipGetErrorString(e);
            }
        }
        if (rc == ISSL_OK) {
            const OutLayout L(s->n);
            const char *col = static_cast<const char *>(s->pin.p) + s->n * 8;
            uint64_t off = 0;
            for (const Piece &pc : s->pieces) {
                const issl_results &o = pc.req->out;
                const uint64_t at = pc.req_off;
                put(o.mit ? o.mit + at : nullptr, col + L.mit, off, pc.n);
                put(o.cfd ? o.cfd + at : nullptr, col + L.cfd, off, pc.n);
                put(o.mit_sum ? o.mit_sum + at : nullptr, col + L.msum, off, pc.n);
                put(o.cfd_sum ? o.cfd_sum + at : nullptr, col + L.csum, off, pc.n);
                put(o.occ_count ? o.occ_count + at : nullptr, col + L.occ, off, pc.n);
                put(o.site_count ? o.site_count + at : nullptr, col + L.sites, off, pc.n);
                put(o.pass ? o.pass + at : nullptr, col + L.pass, off, pc.n);
                off += pc.n;
            }
        }
        lk.lock();
        for (const Piece &pc : s->pieces) {
            const std::shared_ptr<Request> &r = pc.req;
            --r->pending;
            if (rc != ISSL_OK) {
                // the rest of a failed request is not launched
                auto q = std::find(queue.begin(), queue.end(), r);
                if (q != queue.end()) queue.erase(q);
                fail_request(r, rc, s->error);
            } else {
                r->returned += pc.n;
                settle(r);
            }
        }
        s->pieces.clear();
        if (--index_inflight[s->index_id] == 0) index_inflight.erase(s->index_id);
        inflight.pop_front();
        free_slots.push_back(s);
        cv_launch.notify_one();
        cv_done.notify_all();
    }
}

extern "C" int issl_server_create(const issl_server_config *cfg, issl_server **out) {
    if (!cfg || !out) return fail(ISSL_ERR_ARG, "null argument");
    *out = nullptr;
    const int ndev = issl_device_count();
    if (ndev <= 0) return fail(ISSL_ERR_DEVICE, "no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(ISSL_ERR_ARG, "device ordinal out of range");
    if (cfg->slots < 0 || cfg->slots > 16) return fail(ISSL_ERR_ARG, "slots must be 0..16");
    std::unique_ptr<issl_server> s(new (std::nothrow) issl_server);
    if (!s) return fail(ISSL_ERR_NOMEM, "host allocation failed");
    s->cfg = *cfg;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 65536;
    if (s->cfg.slots == 0) s->cfg.slots = 3;
    ISSL_HIP(hipSetDevice(s->cfg.device));
    for (int i = 0; i < s->cfg.slots; ++i) {
        auto sl = std::make_unique<Slot>();
        ISSL_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
        ISSL_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
        s->free_slots.push_back(sl.get());
        s->slots.push_back(std::move(sl));
    }
    issl_server *p = s.get();
    s->launcher = std::thread([p] { p->launch_loop(); });
    s->completer = std::thread([p] { p->complete_loop(); });
    *out = s.release();
    return ok();
}

extern "C" int issl_server_destroy(issl_server *srv) {
    if (!srv) return fail(ISSL_ERR_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (const auto &r : srv->queue)
            srv->fail_request(r, ISSL_ERR_STATE, "server destroyed before the request ran");
        srv->queue.clear();
        srv->stop = true;
    }
    srv->cv_launch.notify_all();
    srv->cv_retire.notify_all();
    if (srv->launcher.joinable()) srv->launcher.join();
    srv->cv_retire.notify_all();
    if (srv->completer.joinable()) srv->completer.join();
    srv->cv_done.notify_all();
    for (auto &sl : srv->slots) {
        if (sl->stream) (void)hipStreamSynchronize(sl->stream);
        sl->guides.release();
        sl->out.release();
        sl->pin.release();
        sl->ws.guides.release();
        sl->ws.out.release();
        sl->ws.tiles.release();
        sl->ws.scratch.release();
        sl->ws.pinned.release();
        if (sl->done) (void)hipEventDestroy(sl->done);
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
    }
    for (auto &kv : srv->indexes) issl_index_close(kv.second);
    delete srv;
    return ok();
}

namespace {

// Admission of an index: free HBM must cover it plus the reserve.
int admit(issl_server *srv, uint64_t need) {
    uint64_t fr = 0, tot = 0;
    int rc = issl_device_memory(srv->cfg.device, &fr, &tot);
    if (rc != ISSL_OK) return rc;
    if (need + srv->cfg.hbm_reserve > fr)
        return fail(ISSL_ERR_NOMEM, "admission refused: index needs " + std::to_string(need) +
                                        " B + reserve " + std::to_string(srv->cfg.hbm_reserve) +
                                        " B, device has " + std::to_string(fr) + " B free");
    return ISSL_OK;
}

int register_index(issl_server *srv, issl_index *ix, int32_t *index_id) {
    issl_index_info info{};
    issl_index_get_info(ix, &info);
    std::lock_guard<std::mutex> lk(srv->mu);
    const int32_t id = srv->next_index++;
    srv->indexes[id] = ix;
    srv->index_bytes[id] = info.device_bytes;
    *index_id = id;
    return ok();
}

}  // namespace

extern "C" int issl_server_add_index(issl_server *srv, const char *path, int32_t *index_id) {
    if (!srv || !path || !index_id) return fail(ISSL_ERR_ARG, "null argument");
    uint64_t n = 0, L = 0, S = 0, nl = 0;
    int rc = issl_read_header(path, &n, &L, &S, &nl);
    if (rc != ISSL_OK) return rc;
    if ((rc = admit(srv, issl_index_bytes_needed(n, S, nl))) != ISSL_OK) return rc;
    issl_index *ix = nullptr;
    if ((rc = issl_index_open(path, srv->cfg.device, &ix)) != ISSL_OK) return rc;
    return register_index(srv, ix, index_id);
}

extern "C" int issl_server_add_index_arrays(issl_server *srv, uint64_t n_sites, uint64_t seq_len,
                                            uint64_t n_slices, const uint64_t *sigs,
                                            const uint64_t *slice_masks,
                                            const uint64_t *list_sizes, const uint64_t *entries,
                                            int32_t *index_id) {
    if (!srv || !index_id) return fail(ISSL_ERR_ARG, "null argument");
    uint64_t n_lists = 0;
    for (uint64_t s = 0; s < n_slices && slice_masks; ++s) {
        uint8_t pos[64];
        const int w = mask_positions(slice_masks[s], seq_len, pos);
        if (w > kMaxWidth) return fail(ISSL_ERR_FORMAT, "slice wider than 16 positions");
        n_lists += 1ULL << (2 * w);
    }
    int rc = admit(srv, issl_index_bytes_needed(n_sites, n_slices, n_lists));
    if (rc != ISSL_OK) return rc;
    issl_index *ix = nullptr;
    if ((rc = issl_index_from_arrays(n_sites, seq_len, n_slices, sigs, slice_masks, list_sizes,
                                     entries, srv->cfg.device, &ix)) != ISSL_OK)
        return rc;
    return register_index(srv, ix, index_id);
}

extern "C" int issl_server_remove_index(issl_server *srv, int32_t index_id) {
    if (!srv) return fail(ISSL_ERR_ARG, "null server");
    issl_index *ix = nullptr;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->indexes.find(index_id);
        if (it == srv->indexes.end()) return fail(ISSL_ERR_ARG, "unknown index id");
        for (auto q = srv->queue.begin(); q != srv->queue.end();) {
            if ((*q)->index_id == index_id) {
                srv->fail_request(*q, ISSL_ERR_STATE, "index was removed");
                q = srv->queue.erase(q);
            } else {
                ++q;
            }
        }
        srv->cv_done.notify_all();
        srv->cv_done.wait(lk, [&] { return srv->index_inflight.count(index_id) == 0; });
        ix = it->second;
        srv->indexes.erase(it);
        srv->index_bytes.erase(index_id);
    }
    return issl_index_close(ix);
}

extern "C" int issl_server_submit(issl_server *srv, int32_t index_id, const uint64_t *guides,
                                  uint64_t n, const issl_params *params, const issl_results *out,
                                  uint64_t *ticket) {
    if (!srv || !params || !out || !ticket) return fail(ISSL_ERR_ARG, "null argument");
    if (n && !guides) return fail(ISSL_ERR_ARG, "null guides");
    auto r = std::make_shared<Request>();
    r->index_id = index_id;
    r->guides.assign(guides, guides + n);
    r->params = *params;
    r->out = *out;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(ISSL_ERR_STATE, "server is shutting down");
    auto it = srv->indexes.find(index_id);
    if (it == srv->indexes.end()) return fail(ISSL_ERR_ARG, "unknown index id");
    const int rc = params_to_scalars(it->second, params, &r->sc);
    if (rc != ISSL_OK) return rc;
    r->ticket = srv->next_ticket++;
    r->t_submit = Clock::now();
    srv->tickets[r->ticket] = r;
    *ticket = r->ticket;
    if (n == 0) {
        srv->settle(r);
        srv->cv_done.notify_all();
        return ok();
    }
    srv->queue.push_back(r);
    srv->cv_launch.notify_one();
    return ok();
}

namespace {

int release(issl_server *srv, uint64_t ticket, int32_t *status) {
    auto it = srv->tickets.find(ticket);
    const std::shared_ptr<Request> r = it->second;
    srv->tickets.erase(it);
    *status = r->status;
    if (r->status != ISSL_OK) return fail(ISSL_OK, r->error);  // message for the caller
    return ok();
}

}  // namespace

extern "C" int issl_server_wait(issl_server *srv, uint64_t ticket, double timeout_s,
                                int32_t *status) {
    if (!srv || !status) return fail(ISSL_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(ISSL_ERR_STATE, "unknown ticket");
    const std::shared_ptr<Request> r = it->second;
    if (timeout_s < 0) {
        srv->cv_done.wait(lk, [&] { return r->done; });
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                      [&] { return r->done; })) {
        return fail(ISSL_ERR_STATE, "request still running");
    }
    return release(srv, ticket, status);
}

extern "C" int issl_server_poll(issl_server *srv, uint64_t ticket, int32_t *done, int32_t *status) {
    if (!srv || !done || !status) return fail(ISSL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(ISSL_ERR_STATE, "unknown ticket");
    *done = it->second->done ? 1 : 0;
    if (!*done) return ok();
    return release(srv, ticket, status);
}

extern "C" int issl_server_get_stats(issl_server *srv, int reset, issl_server_stats *out) {
    if (!srv || !out) return fail(ISSL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = issl_server_stats{};
    out->requests_done = srv->done;
    out->requests_failed = srv->failed;
    out->guides_done = srv->guides;
    out->launches = srv->launches;
    out->queued = srv->queue.size();
    out->peak_inflight = srv->peak_inflight;
    for (const auto &kv : srv->index_bytes) out->resident_bytes += kv.second;
    issl_device_memory(srv->cfg.device, &out->hbm_free, &out->hbm_total);
    if (!srv->lat_ms.empty()) {
        std::vector<double> v = srv->lat_ms;
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto pct = [&](double q) {
            const size_t k = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size())));
            return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        out->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / static_cast<double>(v.size());
    }
    if (srv->launches) out->mean_batch_guides = static_cast<double>(srv->guides) / srv->launches;
    if (reset) {
        srv->lat_ms.clear();
        srv->done = srv->failed = srv->guides = srv->launches = srv->peak_inflight = 0;
    }
    return ok();
}
