// server.cpp -- request loop over resident indexes: admission by free HBM,
// continuous batching of queued requests, per-request latency.
//
// Host code over the public C-ABI only.  One worker thread per server takes
// the oldest queued request, merges every queued request for the same index
// and the same parameters (up to max_batch guides) into one issl_score call,
// and scatters the results.  Each guide's result depends only on the guide
// (batch composition is invisible), so merging never changes a result.
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

#include "common.hpp"
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
    issl_results out{};
    Clock::time_point t_submit;
    bool done = false;
    int32_t status = ISSL_OK;
    std::string error;
};

bool same_params(const issl_params &a, const issl_params &b) {
    return a.max_dist == b.max_dist && a.method == b.method && a.threshold == b.threshold &&
           a.form == b.form && a.precision == b.precision && a.flags == b.flags &&
           a.chunk == b.chunk;
}

// Result columns of a merged launch.
struct Columns {
    std::vector<double> mit, cfd, mit_sum, cfd_sum;
    std::vector<uint64_t> occ, sites;
    std::vector<char> pass;
    void resize(uint64_t n) {
        mit.resize(n);
        cfd.resize(n);
        mit_sum.resize(n);
        cfd_sum.resize(n);
        occ.resize(n);
        sites.resize(n);
        pass.resize(n);
    }
    issl_results view() {
        return issl_results{mit.data(), cfd.data(), occ.data(), sites.data(), pass.data(),
                            mit_sum.data(), cfd_sum.data()};
    }
};

template <typename T>
void put(T *dst, const std::vector<T> &src, uint64_t off, uint64_t n) {
    if (dst) std::memcpy(dst, src.data() + off, n * sizeof(T));
}

}  // namespace

struct issl_server {
    issl_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
    std::map<int32_t, issl_index *> indexes;
    std::map<int32_t, uint64_t> index_bytes;
    int32_t busy_index = -1;       // index of the launch in flight
    uint64_t next_ticket = 1;
    int32_t next_index = 0;
    bool stop = false;
    std::thread worker;
    // statistics since the last reset
    std::vector<double> lat_ms;
    uint64_t done = 0, failed = 0, guides = 0, launches = 0;

    void run();
    void finish(const std::vector<std::shared_ptr<Request>> &group, int rc, const std::string &msg,
                Columns *cols);
};

void issl_server::finish(const std::vector<std::shared_ptr<Request>> &group, int rc,
                         const std::string &msg, Columns *cols) {
    const Clock::time_point now = Clock::now();
    uint64_t off = 0;
    for (const auto &r : group) {
        const uint64_t n = r->guides.size();
        if (rc == ISSL_OK && cols) {
            put(r->out.mit, cols->mit, off, n);
            put(r->out.cfd, cols->cfd, off, n);
            put(r->out.mit_sum, cols->mit_sum, off, n);
            put(r->out.cfd_sum, cols->cfd_sum, off, n);
            put(r->out.occ_count, cols->occ, off, n);
            put(r->out.site_count, cols->sites, off, n);
            put(r->out.pass, cols->pass, off, n);
        }
        off += n;
        r->status = rc;
        r->error = msg;
        r->done = true;
        ++done;
        if (rc != ISSL_OK) ++failed;
        guides += n;
        lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
    }
}

void issl_server::run() {
    Columns cols;
    std::vector<uint64_t> merged;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) break;  // stop requested and nothing left
        // The oldest request and every queued request it can share a launch with.
        std::vector<std::shared_ptr<Request>> group;
        const std::shared_ptr<Request> head = queue.front();
        uint64_t total = 0;
        for (auto it = queue.begin(); it != queue.end();) {
            const auto &r = *it;
            const bool fits = group.empty() || total + r->guides.size() <= cfg.max_batch;
            if (r->index_id == head->index_id && same_params(r->params, head->params) && fits) {
                total += r->guides.size();
                group.push_back(r);
                it = queue.erase(it);
                if (total >= cfg.max_batch) break;
            } else {
                ++it;
            }
        }
        auto ix_it = indexes.find(head->index_id);
        if (ix_it == indexes.end()) {
            finish(group, ISSL_ERR_STATE, "index was removed", nullptr);
            cv_done.notify_all();
            continue;
        }
        issl_index *ix = ix_it->second;
        busy_index = head->index_id;
        ++launches;
        lk.unlock();

        merged.resize(total);
        uint64_t off = 0;
        for (const auto &r : group) {
            std::copy(r->guides.begin(), r->guides.end(), merged.begin() + off);
            off += r->guides.size();
        }
        cols.resize(total);
        issl_results res = cols.view();
        const int rc = issl_score(ix, merged.data(), total, &head->params, &res);
        const std::string msg = rc == ISSL_OK ? std::string() : std::string(issl_last_error());

        lk.lock();
        busy_index = -1;
        finish(group, rc, msg, &cols);
        cv_done.notify_all();
    }
}

extern "C" int issl_server_create(const issl_server_config *cfg, issl_server **out) {
    if (!cfg || !out) return fail(ISSL_ERR_ARG, "null argument");
    *out = nullptr;
    const int ndev = issl_device_count();
    if (ndev <= 0) return fail(ISSL_ERR_DEVICE, "no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(ISSL_ERR_ARG, "device ordinal out of range");
    std::unique_ptr<issl_server> s(new (std::nothrow) issl_server);
    if (!s) return fail(ISSL_ERR_NOMEM, "host allocation failed");
    s->cfg = *cfg;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 65536;
    s->worker = std::thread([p = s.get()] { p->run(); });
    *out = s.release();
    return ok();
}

extern "C" int issl_server_destroy(issl_server *srv) {
    if (!srv) return fail(ISSL_ERR_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        std::vector<std::shared_ptr<Request>> left(srv->queue.begin(), srv->queue.end());
        srv->queue.clear();
        srv->finish(left, ISSL_ERR_STATE, "server destroyed before the request ran", nullptr);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    srv->cv_done.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    for (auto &kv : srv->indexes) issl_index_close(kv.second);
    delete srv;
    return ok();
}

namespace {

// Admission: free HBM must cover the index plus the reserve.
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
        std::vector<std::shared_ptr<Request>> drop;
        for (auto q = srv->queue.begin(); q != srv->queue.end();) {
            if ((*q)->index_id == index_id) {
                drop.push_back(*q);
                q = srv->queue.erase(q);
            } else {
                ++q;
            }
        }
        srv->finish(drop, ISSL_ERR_STATE, "index was removed", nullptr);
        srv->cv_done.notify_all();
        srv->cv_done.wait(lk, [&] { return srv->busy_index != index_id; });
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
    // Parameter check against the index (a zero-guide call scores nothing).
    issl_results none{};
    const int rc = issl_score(it->second, nullptr, 0, params, &none);
    if (rc != ISSL_OK) return rc;
    r->ticket = srv->next_ticket++;
    r->t_submit = Clock::now();
    srv->tickets[r->ticket] = r;
    srv->queue.push_back(r);
    *ticket = r->ticket;
    srv->cv_work.notify_one();
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
    for (const auto &kv : srv->index_bytes) out->resident_bytes += kv.second;
    issl_device_memory(srv->cfg.device, &out->hbm_free, &out->hbm_total);
    if (!srv->lat_ms.empty()) {
        std::vector<double> v = srv->lat_ms;
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto pct = [&](double p) {
            size_t k = static_cast<size_t>(std::ceil(p * v.size()));
            return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        out->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / v.size();
    }
    if (srv->launches) out->mean_batch_guides = static_cast<double>(srv->guides) / srv->launches;
    if (reset) {
        srv->lat_ms.clear();
        srv->done = srv->failed = srv->guides = srv->launches = 0;
    }
    return ok();
}
