// tws_server.cpp -- continuous-batching request loop over a tws_batch.
//
// One thread owns the batch.  Each iteration it
//   1. retires cancelled requests and admits queued ones in arrival order
//      while their footprint fits the pool (a request that does not fit may
//      be passed by later, smaller ones, but only kMaxBypass times; after
//      that admission waits for it);
//   2. launches one batched run that advances every resident request by
//      min(remaining, chunk) steps (asynchronous on the batch stream);
//   3. while that run executes, on a separate I/O stream: downloads the
//      requests that finished in the previous run and uploads the inputs of
//      the requests just admitted (neither is part of the running launch);
//   4. waits for the run, then moves finished requests to the download list.
// Inputs are copied at submit into pinned host buffers (a reusable pool), so
// uploads and downloads are DMA transfers that overlap the kernels.  Because
// each grid's arithmetic does not depend on what else is in the batch,
// results equal the same solve run alone.

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
#include <new>
#include <thread>
#include <vector>

#include "tws.h"
#include "tws_grid.hpp"
#include "tws_runtime.hpp"

using namespace tws;

namespace {

constexpr int kMaxBypass = 8;
constexpr size_t kPinnedCacheBytes = size_t(8) << 30;

// Pinned host buffers, reused by size (hipHostMalloc is slow; the same sizes
// recur).  Thread-safe: the submitter acquires, the server thread releases.
class PinnedPool {
public:
    ~PinnedPool() {
        for (auto& kv : free_) (void)hipHostFree(kv.second);
    }
    void* acquire(size_t bytes) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 2 * bytes) {
                void* p = it->second;
                cached_ -= it->first;
                sizes_[p] = it->first;
                free_.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        sizes_[p] = bytes;
        return p;
    }
    void release(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        const size_t n = sizes_[p];
        sizes_.erase(p);
        if (cached_ + n > kPinnedCacheBytes) {
            (void)hipHostFree(p);
            return;
        }
        free_.emplace(n, p);
        cached_ += n;
    }

private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> sizes_;
    size_t cached_ = 0;
};

struct Req {
    uint64_t ticket = 0;
    tws_request r{};
    float* in = nullptr;       // pinned: terrain, then depth (if given)
    float* flux_in = nullptr;  // pinned: initial flux (if given)
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
    hipStream_t io = nullptr;
    hipEvent_t cleared = nullptr;  // new grids' clears done on the batch stream
    PinnedPool pinned;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    bool broken = false;            // the loop stopped on a host allocation failure
    uint64_t next_ticket = 1;
    std::map<uint64_t, ReqPtr> all;
    std::deque<ReqPtr> queue;
    std::vector<ReqPtr> resident;   // in the batch and advancing
    std::vector<ReqPtr> finishing;  // all steps done, outputs not yet copied
    tws_server_stats stats{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void fail_all() noexcept;
    void finish(const ReqPtr& q, int state, int error);
    int deliver(const ReqPtr& q);
    void free_inputs(const ReqPtr& q) {
        pinned.release(q->in);
        pinned.release(q->flux_in);
        q->in = q->flux_in = nullptr;
    }
};

// Caller holds mu.
void tws_server::finish(const ReqPtr& q, int state, int error) {
    q->st.state = state;
    q->st.error = error;
    q->st.t_done = now();
    const int64_t cu = (int64_t)q->r.width * q->r.height * q->st.steps_done;
    if (state == TWS_REQ_DONE) {
        ++stats.completed;
        stats.cell_updates += cu;
    } else if (state == TWS_REQ_FAILED) {
        ++stats.failed;
        if (error == TWS_ERR_NONFINITE) stats.cell_updates += cu;
    } else if (state == TWS_REQ_CANCELLED) {
        ++stats.cancelled;
    }
    free_inputs(q);
    cv_done.notify_all();
}

// The loop's host bookkeeping threw (out of memory): every request not yet
// finished fails, wherever the loop had it, and the server takes no more.
void tws_server::fail_all() noexcept {
    (void)hipStreamSynchronize(io);
    (void)tws_batch_sync(batch);
    std::lock_guard<std::mutex> lk(mu);
    broken = true;
    try {
        for (auto& kv : all) {
            const ReqPtr& q = kv.second;
            if (q->st.state != TWS_REQ_QUEUED && q->st.state != TWS_REQ_RUNNING) continue;
            if (q->gid >= 0) (void)tws_batch_remove(batch, q->gid);
            finish(q, TWS_REQ_FAILED, TWS_ERR_NOMEM);
        }
        queue.clear();
        resident.clear();
        finishing.clear();
    } catch (...) {
    }
    cv_done.notify_all();
}

// Copies a finished request's outputs to the caller (I/O stream; the grid is
// not part of the running launch sequence).
int tws_server::deliver(const ReqPtr& q) {
    GridState* g = batch_grid(batch, q->gid);
    if (!g) return fail(TWS_ERR_STATE, "server: lost grid %d", q->gid);
    const size_t cells = (size_t)q->r.width * q->r.height;
    int rc = TWS_OK;
    if (q->r.out_depth) {
        float* st = static_cast<float*>(pinned.acquire(cells * sizeof(float)));
        if (!st) return fail(TWS_ERR_NOMEM, "server: pinned staging");
        rc = grid_download_depth_async(*g, st, io);
        if (rc == TWS_OK) {
            hipError_t e = hipStreamSynchronize(io);
            if (e != hipSuccess) rc = hip_fail(e, "server: download");
        }
        if (rc == TWS_OK) std::memcpy(q->r.out_depth, st, cells * sizeof(float));
        pinned.release(st);
    }
    if (rc == TWS_OK && (q->r.out_flux4 || q->r.out_velocity2))
        rc = grid_download(*g, nullptr, q->r.out_flux4, q->r.out_velocity2, io);
    if (rc == TWS_OK) rc = grid_check_finite(*g, io);
    return rc;
}

void tws_server::loop() {
    for (;;) {
        std::vector<ReqPtr> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                return stop || !queue.empty() || !resident.empty() || !finishing.empty();
            });
            if (stop) {
                (void)hipStreamSynchronize(io);
                tws_batch_sync(batch);
                for (auto* list : {&resident, &finishing}) {
                    for (auto& q : *list) {
                        tws_batch_remove(batch, q->gid);
                        finish(q, TWS_REQ_CANCELLED, 0);
                    }
                    list->clear();
                }
                for (auto& q : queue) finish(q, TWS_REQ_CANCELLED, 0);
                queue.clear();
                return;
            }
            for (size_t i = 0; i < resident.size();) {
                if (resident[i]->cancel) {
                    tws_batch_remove(batch, resident[i]->gid);
                    finish(resident[i], TWS_REQ_CANCELLED, 0);
                    resident.erase(resident.begin() + (long)i);
                } else {
                    ++i;
                }
            }
            // admission in arrival order by footprint
            tws_batch_stats bs{};
            tws_batch_stats_get(batch, &bs);
            int slots = cfg.max_resident - (int)(resident.size() + finishing.size());
            bool head_blocked = false;
            for (auto it = queue.begin(); it != queue.end() && slots > 0;) {
                ReqPtr q = *it;
                if (q->footprint <= bs.pool_largest_free) {
                    int gid = -1;
                    if (tws_batch_add(batch, q->r.width, q->r.height, &q->r.params, &gid) == TWS_OK) {
                        q->gid = gid;
                        // RUNNING from here on: a cancel now only flags the request,
                        // which the loop retires once its upload is done
                        q->st.state = TWS_REQ_RUNNING;
                        q->st.t_admit = now();
                        admit.push_back(q);
                        it = queue.erase(it);
                        --slots;
                        tws_batch_stats_get(batch, &bs);
                        continue;
                    }
                }
                if (++q->bypassed > kMaxBypass) head_blocked = true;
                if (head_blocked) break;
                ++it;
            }
            stats.queued = (int64_t)queue.size();
        }
        // the new grids' clears (batch stream) precede their uploads (I/O stream)
        if (!admit.empty()) {
            (void)hipEventRecord(cleared, batch_stream(batch));
            (void)hipStreamWaitEvent(io, cleared, 0);
        }
        // 2. launch one batched run over the resident requests
        std::vector<ReqPtr> run_set = resident;
        std::vector<int> moved(run_set.size(), 0);
        for (size_t i = 0; i < run_set.size(); ++i) {
            const int n = run_set[i]->steps_left < cfg.chunk_steps ? run_set[i]->steps_left
                                                                    : cfg.chunk_steps;
            moved[i] = n;
            tws_batch_set_steps(batch, run_set[i]->gid, n);
        }
        const double t_run = now();
        int rc = run_set.empty() ? TWS_OK : tws_batch_run(batch);
        // 3. overlapped I/O: outputs of the previous run's finishers, inputs of
        //    the admitted requests
        std::vector<ReqPtr> done_now;
        done_now.swap(finishing);
        for (auto& q : done_now) {
            const int drc = deliver(q);
            std::lock_guard<std::mutex> lk(mu);
            tws_batch_remove(batch, q->gid);
            finish(q, drc == TWS_OK ? TWS_REQ_DONE : TWS_REQ_FAILED, drc);
        }
        for (auto& q : admit) {
            GridState* g = batch_grid(batch, q->gid);
            const size_t cells = (size_t)q->r.width * q->r.height;
            int urc = g ? TWS_OK : fail(TWS_ERR_STATE, "server: lost grid");
            if (urc == TWS_OK)
                urc = grid_upload_async(*g, q->in, q->r.depth ? q->in + cells : nullptr, io);
            if (urc == TWS_OK && q->flux_in) {
                hipError_t e = hipStreamSynchronize(io);
                urc = e == hipSuccess ? grid_upload(*g, nullptr, nullptr, q->flux_in, io)
                                      : hip_fail(e, "server: upload");
            }
            if (urc == TWS_OK && q->r.rain_on)
                urc = grid_set_rain(*g, q->r.rain[0], q->r.rain[1], q->r.rain[2], q->r.rain[3]);
            q->st.error = urc;
        }
        hipError_t eio = hipStreamSynchronize(io);
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& q : admit) {
                const int urc = eio != hipSuccess ? TWS_ERR_DEVICE : q->st.error;
                if (urc != TWS_OK) {
                    tws_batch_remove(batch, q->gid);
                    finish(q, TWS_REQ_FAILED, urc);
                    continue;
                }
                free_inputs(q);
                resident.push_back(q);
            }
        }
        // 4. wait for the run
        if (rc == TWS_OK && !run_set.empty()) rc = tws_batch_sync(batch);
        const double t_end = now();
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < run_set.size(); ++i) {
            ReqPtr q = run_set[i];
            auto pos = std::find(resident.begin(), resident.end(), q);
            if (pos == resident.end()) continue;
            if (rc != TWS_OK) {
                resident.erase(pos);
                tws_batch_remove(batch, q->gid);
                finish(q, TWS_REQ_FAILED, rc);
                continue;
            }
            q->steps_left -= moved[i];
            q->st.steps_done += moved[i];
            if (q->steps_left <= 0) {
                resident.erase(pos);
                finishing.push_back(q);
            }
        }
        if (!run_set.empty()) {
            ++stats.runs;
            stats.busy_seconds += t_end - t_run;
        }
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
    // 48: a whole number of launches of the batch default (6 steps per
    // streaming launch, 4 or 8 per tile launch), so a chunk ends in no tail
    // launch, and short enough that arrivals join soon.  MI355X, the served
    // mix at 70% load, p50: 0.136-0.140 s at 48 steps, 0.142 s at 96, 0.151 s
    // at 192; saturated 4.40e11 cell-updates/s at 48 against 4.36-4.38e11 at
    // 96 and 4.28-4.29e11 at 100 (which ends in a tail launch).
    if (c.chunk_steps <= 0) c.chunk_steps = 48;
    if (c.max_queue <= 0) c.max_queue = 1 << 16;
    if (c.pool_bytes <= 0) return fail(TWS_ERR_INVALID, "tws_server_create: pool_bytes must be > 0");
    tws_server* s = new (std::nothrow) tws_server();
    if (!s) return fail(TWS_ERR_NOMEM, "tws_server_create: host allocation failed");
    s->cfg = c;
    int rc = tws_batch_create(c.max_resident, c.pool_bytes, nullptr, &s->batch);
    if (rc == TWS_OK && (c.kernel != TWS_KERNEL_AUTO || c.steps_per_launch != 0))
        rc = tws_batch_set_launch(s->batch, c.kernel, c.steps_per_launch);
    if (rc == TWS_OK) {
        hipError_t e = hipStreamCreateWithFlags(&s->io, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->cleared, hipEventDisableTiming);
        if (e != hipSuccess) rc = hip_fail(e, "tws_server_create: stream");
    }
    if (rc != TWS_OK) {
        if (s->cleared) (void)hipEventDestroy(s->cleared);
        if (s->io) (void)hipStreamDestroy(s->io);
        if (s->batch) tws_batch_destroy(s->batch);
        delete s;
        return rc;
    }
    tws_batch_stats bs{};
    tws_batch_stats_get(s->batch, &bs);
    s->stats.pool_bytes = bs.pool_bytes;
    try {
        s->th = std::thread([s] {
            try {
                s->loop();
            } catch (...) {
                s->fail_all();
            }
        });
    } catch (...) {
        (void)hipEventDestroy(s->cleared);
        (void)hipStreamDestroy(s->io);
        tws_batch_destroy(s->batch);
        delete s;
        return fail(TWS_ERR_NOMEM, "tws_server_create: cannot start the server thread");
    }
    *out = s;
    return TWS_OK;
}

int tws_server_submit(tws_server* s, const tws_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return fail(TWS_ERR_INVALID, "tws_server_submit: NULL argument");
    if (r->width <= 0 || r->height <= 0 || r->steps < 0 || !r->terrain || !valid_params(&r->params))
        return fail(TWS_ERR_INVALID, "tws_server_submit: invalid request");
    if (!batchable(r->width, r->height, r->params))
        return fail(TWS_ERR_INVALID, "tws_server_submit: a reference-dispatch grid of %d x %d (not "
                    "multiples of 32) steps on the single-grid path only", r->width, r->height);
    if (r->rain_on && (!(r->rain[3] > 0.0f) || !std::isfinite(r->rain[0]) ||
                       !std::isfinite(r->rain[1]) || !std::isfinite(r->rain[2])))
        return fail(TWS_ERR_INVALID, "tws_server_submit: invalid rain");
    const int64_t fp = tws_grid_footprint(r->width, r->height);
    if (fp > s->stats.pool_bytes)
        return fail(TWS_ERR_CAPACITY, "tws_server_submit: footprint %lld exceeds the pool",
                    (long long)fp);
    ReqPtr q;
    try {
        q = std::make_shared<Req>();
    } catch (...) {
        return fail(TWS_ERR_NOMEM, "tws_server_submit: host allocation failed");
    }
    q->r = *r;
    q->footprint = fp;
    q->steps_left = r->steps;
    const size_t cells = (size_t)r->width * (size_t)r->height;
    q->in = static_cast<float*>(s->pinned.acquire(sizeof(float) * cells * (r->depth ? 2 : 1)));
    if (!q->in) return fail(TWS_ERR_NOMEM, "tws_server_submit: pinned input buffer");
    std::memcpy(q->in, r->terrain, sizeof(float) * cells);
    if (r->depth) std::memcpy(q->in + cells, r->depth, sizeof(float) * cells);
    if (r->flux4) {
        q->flux_in = static_cast<float*>(s->pinned.acquire(sizeof(float) * 4 * cells));
        if (!q->flux_in) {
            s->free_inputs(q);
            return fail(TWS_ERR_NOMEM, "tws_server_submit: pinned flux buffer");
        }
        std::memcpy(q->flux_in, r->flux4, sizeof(float) * 4 * cells);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->broken) {
        s->free_inputs(q);
        return fail(TWS_ERR_STATE, "tws_server_submit: the server stopped after a host "
                    "allocation failure");
    }
    if ((int64_t)s->queue.size() >= s->cfg.max_queue) {
        s->free_inputs(q);
        return fail(TWS_ERR_CAPACITY, "tws_server_submit: queue full");
    }
    q->ticket = s->next_ticket;
    q->st.state = TWS_REQ_QUEUED;
    q->st.t_submit = s->now();
    try {
        s->all[q->ticket] = q;
        try {
            s->queue.push_back(q);
        } catch (...) {
            s->all.erase(q->ticket);
            throw;
        }
    } catch (...) {
        s->free_inputs(q);
        return fail(TWS_ERR_NOMEM, "tws_server_submit: host allocation failed");
    }
    ++s->next_ticket;
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

int tws_server_release(tws_server* s, uint64_t ticket) {
    if (!s) return fail(TWS_ERR_INVALID, "tws_server_release: NULL handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(TWS_ERR_INVALID, "tws_server_release: unknown ticket");
    const int st = it->second->st.state;
    if (st == TWS_REQ_QUEUED || st == TWS_REQ_RUNNING)
        return fail(TWS_ERR_STATE, "tws_server_release: request %llu not finished",
                    (unsigned long long)ticket);
    s->all.erase(it);
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
    if (s->cleared) (void)hipEventDestroy(s->cleared);
    if (s->io) (void)hipStreamDestroy(s->io);
    delete s;
    return TWS_OK;
}

}  // extern "C"
