// tws_server.cpp -- continuous-batching request loop over a tws_batch.
//
// One thread owns the batch.  Each iteration it
//   1. retires cancelled requests,
//   2. admits queued requests in arrival order while their footprint fits
//      the pool (a request that does not fit may be passed by later, smaller
//      ones, but only kMaxBypass times; after that admission waits for it),
//   3. uploads the admitted requests' inputs,
//   4. advances every resident request by min(remaining, chunk) steps in one
//      batched run,
//   5. downloads finished requests into the callers' buffers and frees their
//      blocks.
// Requests arriving during a run join at the next iteration.  Because each
// grid's arithmetic does not depend on what else is in the batch, results
// equal the same solve run alone.

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "tws.h"
#include "tws_runtime.hpp"

using namespace tws;

namespace {

constexpr int kMaxBypass = 8;

struct Req {
    uint64_t ticket = 0;
    tws_request r{};
    std::vector<float> terrain, depth, flux;
    tws_request_status st{};
    int gid = -1;
    int steps_left = 0;
    int64_t footprint = 0;
    bool cancel = false;
    int bypassed = 0;
};

using ReqPtr = std::shared_ptr<Req>;

}  // namespace

struct tws_server {
    tws_server_config cfg{};
    tws_batch* batch = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_ticket = 1;
    std::map<uint64_t, ReqPtr> all;
    std::deque<ReqPtr> queue;
    std::vector<ReqPtr> resident;  // touched by the server thread only (under mu)
    tws_server_stats stats{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void finish(const ReqPtr& q, int state, int error);
};

// Caller holds mu.
void tws_server::finish(const ReqPtr& q, int state, int error) {
    q->st.state = state;
    q->st.error = error;
    q->st.t_done = now();
    if (state == TWS_REQ_DONE) {
        ++stats.completed;
        stats.cell_updates += (int64_t)q->r.width * q->r.height * q->st.steps_done;
    } else if (state == TWS_REQ_FAILED) {
        ++stats.failed;
        if (error == TWS_ERR_NONFINITE)
            stats.cell_updates += (int64_t)q->r.width * q->r.height * q->st.steps_done;
    } else if (state == TWS_REQ_CANCELLED) {
        ++stats.cancelled;
    }
    q->terrain.clear(); q->terrain.shrink_to_fit();
    q->depth.clear(); q->depth.shrink_to_fit();
    q->flux.clear(); q->flux.shrink_to_fit();
    cv_done.notify_all();
}

void tws_server::loop() {
    for (;;) {
        std::vector<ReqPtr> admit, to_remove;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !resident.empty(); });
            if (stop) {
                for (auto& q : resident) {
                    tws_batch_remove(batch, q->gid);
                    finish(q, TWS_REQ_CANCELLED, 0);
                }
                resident.clear();
                for (auto& q : queue) finish(q, TWS_REQ_CANCELLED, 0);
                queue.clear();
                return;
            }
            // 1. cancelled residents leave
            for (size_t i = 0; i < resident.size();) {
                if (resident[i]->cancel) {
                    tws_batch_remove(batch, resident[i]->gid);
                    finish(resident[i], TWS_REQ_CANCELLED, 0);
                    resident.erase(resident.begin() + (long)i);
                } else {
                    ++i;
                }
            }
            // 2. admission in arrival order by footprint
            tws_batch_stats bs{};
            tws_batch_stats_get(batch, &bs);
            int64_t free_bytes = bs.pool_bytes - bs.pool_used;
            int64_t largest = bs.pool_largest_free;
            int slots = cfg.max_resident - (int)resident.size();
            bool head_blocked = false;
            for (auto it = queue.begin(); it != queue.end() && slots > 0;) {
                ReqPtr q = *it;
                if (q->footprint <= largest && q->footprint <= free_bytes) {
                    int gid = -1;
                    if (tws_batch_add(batch, q->r.width, q->r.height, &q->r.params, &gid) == TWS_OK) {
                        q->gid = gid;
                        admit.push_back(q);
                        it = queue.erase(it);
                        --slots;
                        tws_batch_stats_get(batch, &bs);
                        free_bytes = bs.pool_bytes - bs.pool_used;
                        largest = bs.pool_largest_free;
                        continue;
                    }
                }
                if (++q->bypassed > kMaxBypass) head_blocked = true;
                if (head_blocked) break;  // wait for this one to fit
                ++it;
            }
            stats.queued = (int64_t)queue.size();
        }
        // 3. upload admitted inputs (outside the lock; the batch is ours)
        for (auto& q : admit) {
            int rc = tws_batch_upload(batch, q->gid, q->terrain.data(),
                                      q->depth.empty() ? nullptr : q->depth.data(),
                                      q->flux.empty() ? nullptr : q->flux.data());
            if (rc == TWS_OK && q->r.rain_on)
                rc = tws_batch_set_rain(batch, q->gid, q->r.rain[0], q->r.rain[1], q->r.rain[2],
                                        q->r.rain[3]);
            std::lock_guard<std::mutex> lk(mu);
            if (rc != TWS_OK) {
                tws_batch_remove(batch, q->gid);
                finish(q, TWS_REQ_FAILED, rc);
                continue;
            }
            q->st.state = TWS_REQ_RUNNING;
            q->st.t_admit = now();
            q->terrain.clear(); q->terrain.shrink_to_fit();
            q->depth.clear(); q->depth.shrink_to_fit();
            q->flux.clear(); q->flux.shrink_to_fit();
            resident.push_back(q);
        }
        std::vector<ReqPtr> run_set;
        {
            std::lock_guard<std::mutex> lk(mu);
            run_set = resident;
            stats.resident = (int64_t)resident.size();
        }
        if (run_set.empty()) continue;
        // 4. one batched run
        std::vector<int> moved(run_set.size(), 0);
        for (size_t i = 0; i < run_set.size(); ++i) {
            const int n = run_set[i]->steps_left < cfg.chunk_steps ? run_set[i]->steps_left
                                                                    : cfg.chunk_steps;
            moved[i] = n;
            tws_batch_set_steps(batch, run_set[i]->gid, n);
        }
        const double t_run = now();
        int rc = tws_batch_run(batch);
        if (rc == TWS_OK) rc = tws_batch_sync(batch);
        const double t_end = now();
        // 5. retire finished requests
        for (size_t i = 0; i < run_set.size(); ++i) {
            ReqPtr q = run_set[i];
            int final_rc = rc;
            bool done = false;
            if (rc == TWS_OK) {
                q->steps_left -= moved[i];
                q->st.steps_done += moved[i];
                done = q->steps_left <= 0;
            }
            if (rc != TWS_OK || done) {
                if (rc == TWS_OK) {
                    final_rc = tws_batch_download(batch, q->gid, q->r.out_depth, q->r.out_flux4,
                                                  q->r.out_velocity2);
                    if (final_rc == TWS_OK) final_rc = tws_batch_check_finite(batch, q->gid);
                }
                std::lock_guard<std::mutex> lk(mu);
                tws_batch_remove(batch, q->gid);
                for (size_t j = 0; j < resident.size(); ++j)
                    if (resident[j] == q) { resident.erase(resident.begin() + (long)j); break; }
                finish(q, final_rc == TWS_OK ? TWS_REQ_DONE : TWS_REQ_FAILED, final_rc);
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        ++stats.runs;
        stats.busy_seconds += t_end - t_run;
        stats.resident = (int64_t)resident.size();
        tws_batch_stats bs{};
        if (tws_batch_stats_get(batch, &bs) == TWS_OK) stats.pool_used = bs.pool_used;
    }
}

extern "C" {

int tws_server_create(const tws_server_config* config, tws_server** out) {
    if (!config || !out) return fail(TWS_ERR_INVALID, "tws_server_create: NULL argument");
    *out = nullptr;
    tws_server_config c = *config;
    if (c.max_resident <= 0) c.max_resident = 64;
    if (c.chunk_steps <= 0) c.chunk_steps = 100;
    if (c.max_queue <= 0) c.max_queue = 1 << 16;
    if (c.pool_bytes <= 0) return fail(TWS_ERR_INVALID, "tws_server_create: pool_bytes must be > 0");
    tws_server* s = new (std::nothrow) tws_server();
    if (!s) return fail(TWS_ERR_NOMEM, "tws_server_create: host allocation failed");
    s->cfg = c;
    int rc = tws_batch_create(c.max_resident, c.pool_bytes, nullptr, &s->batch);
    if (rc == TWS_OK && (c.kernel != TWS_KERNEL_AUTO || c.steps_per_launch != 0))
        rc = tws_batch_set_launch(s->batch, c.kernel, c.steps_per_launch);
    if (rc != TWS_OK) {
        if (s->batch) tws_batch_destroy(s->batch);
        delete s;
        return rc;
    }
    tws_batch_stats bs{};
    tws_batch_stats_get(s->batch, &bs);
    s->stats.pool_bytes = bs.pool_bytes;
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return TWS_OK;
}

int tws_server_submit(tws_server* s, const tws_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return fail(TWS_ERR_INVALID, "tws_server_submit: NULL argument");
    if (r->width <= 0 || r->height <= 0 || r->steps < 0 || !r->terrain || !valid_params(&r->params))
        return fail(TWS_ERR_INVALID, "tws_server_submit: invalid request");
    const int64_t fp = tws_grid_footprint(r->width, r->height);
    if (fp > s->stats.pool_bytes)
        return fail(TWS_ERR_CAPACITY, "tws_server_submit: footprint %lld exceeds the pool",
                    (long long)fp);
    auto q = std::make_shared<Req>();
    q->r = *r;
    q->footprint = fp;
    q->steps_left = r->steps;
    const size_t cells = (size_t)r->width * (size_t)r->height;
    try {
        q->terrain.assign(r->terrain, r->terrain + cells);
        if (r->depth) q->depth.assign(r->depth, r->depth + cells);
        if (r->flux4) q->flux.assign(r->flux4, r->flux4 + 4 * cells);
    } catch (const std::bad_alloc&) {
        return fail(TWS_ERR_NOMEM, "tws_server_submit: host copy failed");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if ((int64_t)s->queue.size() >= s->cfg.max_queue)
        return fail(TWS_ERR_CAPACITY, "tws_server_submit: queue full");
    q->ticket = s->next_ticket++;
    q->st.state = TWS_REQ_QUEUED;
    q->st.t_submit = s->now();
    s->all[q->ticket] = q;
    s->queue.push_back(q);
    ++s->stats.submitted;
    s->stats.queued = (int64_t)s->queue.size();
    *ticket = q->ticket;
    s->cv_work.notify_one();
    return TWS_OK;
}

int tws_server_poll(tws_server* s, uint64_t ticket, tws_request_status* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(TWS_ERR_INVALID, "tws_server_poll: unknown ticket");
    *out = it->second->st;
    return TWS_OK;
}

int tws_server_wait(tws_server* s, uint64_t ticket, double timeout, tws_request_status* out) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_server_wait: NULL handle");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(TWS_ERR_INVALID, "tws_server_wait: unknown ticket");
    ReqPtr q = it->second;
    auto finished = [&] { return q->st.state != TWS_REQ_QUEUED && q->st.state != TWS_REQ_RUNNING; };
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout), finished);
    if (out) *out = q->st;
    if (!ok) return fail(TWS_ERR_STATE, "tws_server_wait: timeout");
    return TWS_OK;
}

int tws_server_cancel(tws_server* s, uint64_t ticket) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_server_cancel: NULL handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(TWS_ERR_INVALID, "tws_server_cancel: unknown ticket");
    ReqPtr q = it->second;
    if (q->st.state == TWS_REQ_QUEUED) {
        for (auto qi = s->queue.begin(); qi != s->queue.end(); ++qi)
            if (*qi == q) { s->queue.erase(qi); break; }
        s->stats.queued = (int64_t)s->queue.size();
        s->finish(q, TWS_REQ_CANCELLED, 0);
    } else if (q->st.state == TWS_REQ_RUNNING) {
        q->cancel = true;
        s->cv_work.notify_one();
    }
    return TWS_OK;
}

int tws_server_drain(tws_server* s) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_server_drain: NULL handle");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return s->stats.completed + s->stats.failed + s->stats.cancelled == s->stats.submitted;
    });
    return TWS_OK;
}

int tws_server_stats_get(tws_server* s, tws_server_stats* out) {
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_server_stats_get: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;  // pool_used as of the last batched run
    return TWS_OK;
}

double tws_server_now(const tws_server* s) { return s ? s->now() : 0.0; }

int tws_server_destroy(tws_server* s) {
    if (!s) return TWS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    tws_batch_destroy(s->batch);
    delete s;
    return TWS_OK;
}

}  // extern "C"
