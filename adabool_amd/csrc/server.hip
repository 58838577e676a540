// server.hip -- serving loop: a resident index, requests submitted from any
// host thread, memory-based admission of result memory, and continuous
// batching into device-dispatched batches on several in-flight streams.
//
// Life of a request
//   submit   keys copied, upper bound of its id count computed on the host
//            (sum of the keys' list lengths, capped at nvalues);
//   admit    that many ids reserved in the device result arena, in arrival
//            order (a request that does not fit waits; later ones wait behind
//            it, so admission is FIFO);
//   batch    admitted requests are packed into the next batch when a stream
//            slot is free and either max_batch requests are waiting, the
//            oldest has waited max_wait_us, or the device is idle;
//   run      one ab_query_batch_packed call per batch on the slot's context:
//            per-query dispatch, one persistent union launch, ids written
//            straight into each request's reservation; counts and paths come
//            back through pinned memory behind one event (ids_only: one
//            ab_query_batch_ids call, no bitmaps);
//   deliver  deliver 0: ids stay in device memory until released;
//            deliver 1: ids are copied to a pinned host mirror of the arena;
//   release  ab_server_release returns the reservation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

using namespace ab;
using Clock = std::chrono::steady_clock;

namespace {

double ms_since(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Result arena of `cap` u32 ids: next-fit over a coalescing free list.  The
// search starts where the last reservation ended, so consecutive requests get
// consecutive ranges (one copy per batch when results go to the host) and
// releases in any order return space at once.
struct Arena {
    uint64_t cap = 0, used = 0, rover = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length

    void init(uint64_t words)
    {
        cap = words;
        used = rover = 0;
        free_.clear();
        free_[0] = words;
    }

    static uint64_t round(uint64_t len) { return (len + 3) & ~3ull; }  // 16-byte aligned ranges

    bool reserve(uint64_t len, uint64_t *off)
    {
        len = round(len);
        if (len == 0) {
            *off = 0;
            return true;
        }
        auto take = [&](std::map<uint64_t, uint64_t>::iterator it) {
            const uint64_t o = it->first, l = it->second;
            free_.erase(it);
            if (l > len) free_[o + len] = l - len;
            used += len;
            rover = o + len;
            *off = o;
            return true;
        };
        for (auto it = free_.lower_bound(rover); it != free_.end(); ++it)
            if (it->second >= len) return take(it);
        for (auto it = free_.begin(); it != free_.end() && it->first < rover; ++it)
            if (it->second >= len) return take(it);
        return false;
    }

    void release(uint64_t off, uint64_t len)
    {
        len = round(len);
        if (len == 0) return;
        used -= len;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    uint64_t ticket;
    std::vector<uint32_t> keys;
    uint64_t ub = 0;   // id upper bound
    uint64_t off = 0;  // reservation (ids) in the arena
    uint32_t count = 0;
    int path = 0;
    int status = AB_OK;
    bool reserved = false;  // holds arena space until released
    bool done = false;
    Clock::time_point t_submit, t_done;
};

struct Slot {
    ab_ctx *ctx = nullptr;
    PinBuf hkeys, hcounts, hpaths;
    DevBuf dkeys, dcounts;
    hipEvent_t ev = nullptr, ev_copy = nullptr;
    std::vector<Request *> members;
    std::vector<uint64_t> qoff, idoff;
    bool busy = false, copying = false;
};

}  // namespace

struct ab_server {
    ab_index *ix = nullptr;
    ab_server_config cfg{};
    std::vector<Slot> slots;
    DevBuf arena;
    PinBuf harena;
    Arena ring;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending, admitted, completed;
    std::unordered_map<uint64_t, Request *> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::string err;
    std::thread worker;
    int device = 0;
    // statistics
    std::vector<double> lat_ms;
    uint64_t batches = 0, batched_queries = 0, dense_queries = 0;
    Clock::time_point t_first, t_last;
    bool any = false;
};

namespace {

// Worker: admission, batch formation, launch, completion.  Holds the lock
// except while it talks to the device.
void serve(ab_server *s)
{
    (void)hipSetDevice(s->device);
    const uint32_t nslots = (uint32_t)s->slots.size();
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // ---- admission (arrival order) ------------------------------------
        while (!s->pending.empty()) {
            Request *r = s->pending.front();
            uint64_t off = 0;
            if (s->ring.reserve(r->ub, &off)) {
                r->off = off;
                r->reserved = true;
                s->admitted.push_back(r);
                s->pending.pop_front();
                continue;
            }
            bool idle = s->admitted.empty();
            for (auto &sl : s->slots) idle = idle && !sl.busy && !sl.copying;
            // never fits, or (shutting down) waits on results nobody will release
            if (idle && (s->ring.used == 0 || s->stop)) {
                r->status = AB_OOM;
                r->done = true;
                r->t_done = Clock::now();
                s->pending.pop_front();
                s->completed.push_back(r);
                s->cv_done.notify_all();
                continue;
            }
            break;
        }
        // ---- completion ------------------------------------------------------
        bool progressed = false;
        for (auto &sl : s->slots) {
            if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
                sl.busy = false;
                progressed = true;
                const uint32_t *cnt = (const uint32_t *)sl.hcounts.p;
                const uint8_t *pth = (const uint8_t *)sl.hpaths.p;
                for (size_t i = 0; i < sl.members.size(); ++i) {
                    sl.members[i]->count = cnt[i];
                    sl.members[i]->path = pth[i];
                    s->dense_queries += pth[i] == AB_PATH_DENSE;
                }
                if (s->cfg.deliver == 1) {
                    // ids -> pinned host mirror: one copy per run of members
                    // whose reservations follow each other with no gap, so a
                    // copy never spans arena space that another request (on
                    // another slot's stream) may still be writing
                    lk.unlock();
                    hipStream_t st = sl.ctx->stream;
                    hipError_t e = hipSuccess;
                    size_t i = 0;
                    while (i < sl.members.size() && e == hipSuccess) {
                        uint64_t a = sl.members[i]->off, b = a + sl.members[i]->count;
                        uint64_t res_end = a + Arena::round(sl.members[i]->ub);  // end of the run's reservations
                        size_t j = i + 1;
                        while (j < sl.members.size() && sl.members[j]->off == res_end) {
                            b = sl.members[j]->off + sl.members[j]->count;
                            res_end = sl.members[j]->off + Arena::round(sl.members[j]->ub);
                            ++j;
                        }
                        if (b > a)
                            e = hipMemcpyAsync((uint32_t *)s->harena.p + a, (const uint32_t *)s->arena.p + a,
                                               (b - a) * 4, hipMemcpyDeviceToHost, st);
                        i = j;
                    }
                    if (e == hipSuccess) e = hipEventRecord(sl.ev_copy, st);
                    lk.lock();
                    if (e != hipSuccess) s->err = std::string("result copy: ") + hipGetErrorString(e);
                    sl.copying = true;
                } else {
                    const auto now = Clock::now();
                    for (Request *r : sl.members) {
                        r->done = true;
                        r->t_done = now;
                        s->completed.push_back(r);
                    }
                    sl.members.clear();
                    s->cv_done.notify_all();
                }
            }
            if (sl.copying && hipEventQuery(sl.ev_copy) == hipSuccess) {
                sl.copying = false;
                progressed = true;
                const auto now = Clock::now();
                for (Request *r : sl.members) {
                    r->done = true;
                    r->t_done = now;
                    s->completed.push_back(r);
                }
                sl.members.clear();
                s->cv_done.notify_all();
            }
        }
        // ---- batch formation and launch --------------------------------------
        Slot *free_slot = nullptr;
        bool any_busy = false;
        for (auto &sl : s->slots) {
            if (!sl.busy && !sl.copying && !free_slot) free_slot = &sl;
            any_busy = any_busy || sl.busy;
        }
        if (free_slot && !s->admitted.empty()) {
            const auto now = Clock::now();
            const double waited = ms_since(s->admitted.front()->t_submit, now) * 1e3;
            const bool full = s->admitted.size() >= s->cfg.max_batch;
            if (full || !any_busy || waited >= (double)s->cfg.max_wait_us || s->stop) {
                Slot &sl = *free_slot;
                const size_t nq = std::min<size_t>(s->admitted.size(), s->cfg.max_batch);
                sl.members.assign(s->admitted.begin(), s->admitted.begin() + nq);
                s->admitted.erase(s->admitted.begin(), s->admitted.begin() + nq);
                sl.busy = true;
                lk.unlock();
                // pack keys and offsets
                sl.qoff.assign(nq + 1, 0);
                sl.idoff.resize(nq);
                for (size_t i = 0; i < nq; ++i) {
                    sl.qoff[i + 1] = sl.qoff[i] + sl.members[i]->keys.size();
                    sl.idoff[i] = sl.members[i]->off;
                }
                const uint64_t K = sl.qoff[nq];
                ab_status st = AB_OK;
                hipError_t e = hipSuccess;
                if ((st = sl.hkeys.ensure(K * 4 + 16)) == AB_OK && (st = sl.dkeys.ensure(K * 4 + 16)) == AB_OK &&
                    (st = sl.dcounts.ensure(nq * 4 + 16)) == AB_OK && (st = sl.hcounts.ensure(nq * 4 + 16)) == AB_OK &&
                    (st = sl.hpaths.ensure(nq + 16)) == AB_OK) {
                    uint32_t *hk = (uint32_t *)sl.hkeys.p;
                    for (size_t i = 0; i < nq; ++i)
                        if (!sl.members[i]->keys.empty())
                            memcpy(hk + sl.qoff[i], sl.members[i]->keys.data(), sl.members[i]->keys.size() * 4);
                    hipStream_t str = sl.ctx->stream;
                    if (K) e = hipMemcpyAsync(sl.dkeys.p, sl.hkeys.p, K * 4, hipMemcpyHostToDevice, str);
                    if (e == hipSuccess && s->cfg.ids_only) {
                        st = ab_query_batch_ids(sl.ctx, (const uint32_t *)sl.dkeys.p, sl.qoff.data(), (uint32_t)nq,
                                                (uint32_t *)s->arena.p, sl.idoff.data(), (uint32_t *)sl.dcounts.p);
                        memset(sl.hpaths.p, AB_PATH_IDS, nq);
                    } else if (e == hipSuccess) {
                        st = ab_query_batch_packed(sl.ctx, (const uint32_t *)sl.dkeys.p, sl.qoff.data(), (uint32_t)nq,
                                                   s->cfg.path, nullptr, (uint32_t *)s->arena.p, sl.idoff.data(),
                                                   (uint32_t *)sl.dcounts.p, nullptr);
                    }
                    if (e == hipSuccess && st == AB_OK)
                        e = hipMemcpyAsync(sl.hcounts.p, sl.dcounts.p, nq * 4, hipMemcpyDeviceToHost, str);
                    if (e == hipSuccess && st == AB_OK && !s->cfg.ids_only)
                        e = hipMemcpyAsync(sl.hpaths.p, sl.ctx->sel.p, nq, hipMemcpyDeviceToHost, str);
                    if (e == hipSuccess && st == AB_OK) e = hipEventRecord(sl.ev, str);
                }
                lk.lock();
                s->batches++;
                s->batched_queries += nq;
                if (st != AB_OK || e != hipSuccess) {
                    s->err = st != AB_OK ? std::string(ab_last_error()) : std::string(hipGetErrorString(e));
                    // fail the batch's requests; the slot's stream is drained before reuse
                    lk.unlock();
                    (void)hipStreamSynchronize(sl.ctx->stream);
                    lk.lock();
                    sl.busy = false;
                    const auto t = Clock::now();
                    for (Request *r : sl.members) {
                        r->status = st != AB_OK ? st : AB_DEVICE_ERROR;
                        r->done = true;
                        r->t_done = t;
                        s->completed.push_back(r);
                    }
                    sl.members.clear();
                    s->cv_done.notify_all();
                }
                progressed = true;
            }
        }
        bool inflight = false;
        for (auto &sl : s->slots) inflight = inflight || sl.busy || sl.copying;
        if (s->stop && s->pending.empty() && s->admitted.empty() && !inflight) break;
        if (progressed) continue;
        // nothing moved: wait for a submission, a release or a device event
        if (inflight || !s->admitted.empty())
            s->cv_work.wait_for(lk, std::chrono::microseconds(inflight ? 10 : 20));
        else
            s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        (void)nslots;
    }
}

}  // namespace

extern "C" ab_status ab_server_config_default(ab_server_config *cfg)
{
    if (!cfg) {
        set_error("ab_server_config_default: NULL argument");
        return AB_INVALID_ARG;
    }
    cfg->max_batch = 1024;
    cfg->max_wait_us = 200;
    cfg->inflight = 2;
    cfg->arena_bytes = 1ull << 32;
    cfg->path = AB_PATH_AUTO;
    cfg->deliver = 0;
    cfg->batch_mode = 1;
    cfg->ids_only = 0;
    return AB_OK;
}

extern "C" ab_status ab_server_create(ab_index *ix, const ab_server_config *cfg, ab_server **out)
{
    if (!ix || !cfg || !out || ix->host_only) {
        set_error("ab_server_create: bad argument (index must be device-resident)");
        return AB_INVALID_ARG;
    }
    if (cfg->max_batch < 1 || cfg->max_batch > 4096 || cfg->inflight < 1 || cfg->inflight > 8 ||
        cfg->arena_bytes < 4096 || cfg->deliver < 0 || cfg->deliver > 1 || cfg->path < 0 ||
        cfg->path > AB_PATH_LEARNED || cfg->batch_mode < 0 || cfg->batch_mode > 2 ||
        cfg->ids_only < 0 || cfg->ids_only > 1) {
        set_error("ab_server_create: config out of range");
        return AB_INVALID_ARG;
    }
    std::unique_ptr<ab_server> s(new ab_server());
    s->ix = ix;
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    s->slots.resize(cfg->inflight);
    for (auto &sl : s->slots) {
        ab_status st = ab_ctx_create(ix, nullptr, &sl.ctx);
        if (st != AB_OK) {
            for (auto &o : s->slots)
                if (o.ctx) ab_ctx_free(o.ctx);
            return st;
        }
        (void)ab_ctx_set_batch_mode(sl.ctx, cfg->batch_mode);
        if (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev_copy, hipEventDisableTiming) != hipSuccess) {
            set_error("ab_server_create: hipEventCreate failed");
            return AB_DEVICE_ERROR;
        }
    }
    const uint64_t words = cfg->arena_bytes / 4;
    ab_status st = s->arena.ensure(words * 4);
    if (st == AB_OK && cfg->deliver == 1) st = s->harena.ensure(words * 4);
    if (st != AB_OK) {
        for (auto &o : s->slots) ab_ctx_free(o.ctx);
        return st;
    }
    s->ring.init(words);
    s->worker = std::thread(serve, s.get());
    *out = s.release();
    return AB_OK;
}

extern "C" ab_ctx *ab_server_ctx(ab_server *s, uint32_t i)
{
    if (!s || i >= s->slots.size()) return nullptr;
    return s->slots[i].ctx;
}

extern "C" ab_status ab_server_submit(ab_server *s, const uint32_t *h_keys, uint32_t n, uint64_t *ticket)
{
    if (!s || (n && !h_keys) || !ticket) {
        set_error("ab_server_submit: bad argument");
        return AB_INVALID_ARG;
    }
    Request *r = new Request();
    r->keys.assign(h_keys, h_keys + n);
    r->ub = id_upper_bound(s->ix, h_keys, n);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) {
            delete r;
            set_error("ab_server_submit: server is shutting down");
            return AB_INVALID_ARG;
        }
        r->ticket = s->next_ticket++;
        s->live[r->ticket] = r;
        s->pending.push_back(r);
        if (!s->any) {
            s->any = true;
            s->t_first = r->t_submit;
        }
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return AB_OK;
}

extern "C" ab_status ab_server_poll(ab_server *s, ab_result *out, uint32_t cap, uint32_t *n, int wait_ms)
{
    if (!s || !n || (cap && !out)) {
        set_error("ab_server_poll: bad argument");
        return AB_INVALID_ARG;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completed.empty() && wait_ms != 0) {
        auto pred = [&] { return !s->completed.empty(); };
        if (wait_ms < 0)
            s->cv_done.wait(lk, pred);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    uint32_t k = 0;
    while (k < cap && !s->completed.empty()) {
        Request *r = s->completed.front();
        s->completed.pop_front();
        ab_result &o = out[k++];
        o.ticket = r->ticket;
        o.status = r->status;
        o.count = r->count;
        o.path = r->path;
        o.ids = r->status != AB_OK ? nullptr
                : s->cfg.deliver == 1 ? (const uint32_t *)s->harena.p + r->off
                                      : (const uint32_t *)s->arena.p + r->off;
        o.latency_ms = ms_since(r->t_submit, r->t_done);
        if (r->status == AB_OK) s->lat_ms.push_back(o.latency_ms);
        if (r->t_done > s->t_last) s->t_last = r->t_done;
    }
    *n = k;
    return s->err.empty() ? AB_OK : (set_error(s->err), AB_DEVICE_ERROR);
}

extern "C" ab_status ab_server_release(ab_server *s, uint64_t ticket)
{
    if (!s) {
        set_error("ab_server_release: NULL server");
        return AB_INVALID_ARG;
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->live.find(ticket);
        if (it == s->live.end() || !it->second->done) {
            set_error("ab_server_release: unknown or unfinished ticket");
            return AB_INVALID_ARG;
        }
        Request *r = it->second;
        if (r->reserved) s->ring.release(r->off, r->ub);
        s->live.erase(it);
        delete r;
    }
    s->cv_work.notify_one();
    return AB_OK;
}

extern "C" ab_status ab_server_stats(ab_server *s, ab_server_stats_t *out, int reset)
{
    if (!s || !out) {
        set_error("ab_server_stats: bad argument");
        return AB_INVALID_ARG;
    }
    std::lock_guard<std::mutex> g(s->mu);
    std::vector<double> l = s->lat_ms;
    double st[9] = {0};
    if (!l.empty()) ab_stats(l.data(), l.size(), st);
    out->completed = s->lat_ms.size();
    out->batches = s->batches;
    out->mean_batch = s->batches ? (double)s->batched_queries / (double)s->batches : 0.0;
    out->dense_queries = s->dense_queries;
    const double span = s->any ? ms_since(s->t_first, s->t_last) * 1e-3 : 0.0;
    out->qps = span > 0 ? (double)s->lat_ms.size() / span : 0.0;
    out->lat_mean_ms = st[0];
    out->lat_p50_ms = st[5];
    out->lat_p90_ms = st[6];
    out->lat_p99_ms = st[8];
    out->lat_max_ms = st[4];
    out->arena_used_bytes = s->ring.used * 4;
    if (reset) {
        s->lat_ms.clear();
        s->batches = s->batched_queries = s->dense_queries = 0;
        s->any = false;
    }
    return AB_OK;
}

extern "C" ab_status ab_server_free(ab_server *s)
{
    if (!s) return AB_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &sl : s->slots) {
        if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.ev_copy) (void)hipEventDestroy(sl.ev_copy);
        sl.hkeys.release();
        sl.hcounts.release();
        sl.hpaths.release();
        sl.dkeys.release();
        sl.dcounts.release();
        if (sl.ctx) ab_ctx_free(sl.ctx);
    }
    for (auto &kv : s->live) delete kv.second;
    s->arena.release();
    s->harena.release();
    delete s;
    return AB_OK;
}

// Open-loop replay: queries arrive at Poisson times of rate `rate_qps`
// (rate 0: all at once) from a submitting thread while this thread collects
// and releases results.  Latency is measured from each query's scheduled
// arrival, so a submitter that falls behind is charged to the server.
extern "C" ab_status ab_server_replay(ab_server *s, const uint32_t *h_keys, const uint64_t *h_qoff, uint32_t nq,
                                      double rate_qps, uint64_t seed, uint32_t *h_counts, double *h_lat_ms,
                                      double *wall_s)
{
    if (!s || !h_qoff || (h_qoff[nq] && !h_keys) || rate_qps < 0) {
        set_error("ab_server_replay: bad argument");
        return AB_INVALID_ARG;
    }
    std::vector<Clock::time_point> when(nq);
    std::vector<uint64_t> tickets(nq, 0);
    std::unordered_map<uint64_t, uint32_t> index;
    std::mutex imu;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_qps > 0 ? rate_qps : 1.0);
    const auto t0 = Clock::now() + std::chrono::milliseconds(2);
    double t = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        if (rate_qps > 0) t += gap(rng);
        when[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t));
    }
    std::atomic<int> sub_status{AB_OK};
    std::vector<Clock::time_point> sub_at(nq);
    std::thread submitter([&] {
        for (uint32_t i = 0; i < nq; ++i) {
            while (Clock::now() < when[i]) {
            }
            sub_at[i] = Clock::now();
            uint64_t tk = 0;
            ab_status st = ab_server_submit(s, h_keys + h_qoff[i], (uint32_t)(h_qoff[i + 1] - h_qoff[i]), &tk);
            if (st != AB_OK) {
                sub_status = st;
                return;
            }
            std::lock_guard<std::mutex> g(imu);
            index[tk] = i;
        }
    });
    uint32_t got = 0;
    std::vector<ab_result> res(256);
    ab_status status = AB_OK;
    while (got < nq && sub_status == AB_OK) {
        uint32_t n = 0;
        ab_status st = ab_server_poll(s, res.data(), (uint32_t)res.size(), &n, 50);
        if (st != AB_OK) status = st;
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t i;
            for (;;) {  // the submitter records the ticket just after submit returns
                std::lock_guard<std::mutex> g(imu);
                auto it = index.find(res[j].ticket);
                if (it != index.end()) {
                    i = it->second;
                    break;
                }
            }
            if (res[j].status != AB_OK) status = (ab_status)res[j].status;
            if (h_counts) h_counts[i] = res[j].count;
            // submit -> done, plus how late the submitter was against the schedule
            if (h_lat_ms) h_lat_ms[i] = res[j].latency_ms + ms_since(when[i], sub_at[i]);
            (void)ab_server_release(s, res[j].ticket);
            ++got;
        }
        if (status != AB_OK) break;
    }
    submitter.join();
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    if (sub_status != AB_OK) return (ab_status)sub_status.load();
    return status;
}
