// qvti_serve.hip -- continuous-batching server for shot-modelling requests
// of uneven grid size and duration (no counterpart in the reference, whose
// host runs shots one after another, viscoelas_3D.cpp:355-438).
//
// A worker thread owns the GPU.  Requests wait in a FIFO until the HBM budget
// and the active-slot limit admit them; an admitted request gets an engine
// context (its own HIP stream) whose wavefields, C-PML memory variables and
// records stay resident until the shot ends.  Every scheduling round enqueues
// a slice of steps on each active request's stream (at most two slices in
// flight per request), so kernels of different requests overlap on the
// device; requests that arrive meanwhile are admitted as soon as a slot and
// memory free up.  Finished contexts return to a pool keyed by grid, so a
// later request of the same shape reuses them without hipMalloc/hipFree.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "qvti.h"
#include "qvti_internal.h"

// The HIP stream a context enqueues its work on (defined in qvti_engine.hip).
hipStream_t qvti_ctx_stream(qvti_ctx* c);

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    qvti_grid grid{};
    int nt = 0;
    std::vector<float> model[7];  // vp, gp, vs, gs, rho, eps, delta
    std::vector<float> rik;
    int kind = 0, sx = 0, sy = 0, sz = 0;
    std::vector<int> rx, ry, rz;
    std::vector<float> rec;  // [3][nt][nrec]
    int status = QVTI_E_PENDING;
    bool cancel = false;
    double t_submit = 0, t_admit = -1, t_done = -1;
    // while active
    qvti_ctx* ctx = nullptr;
    long long bytes = 0;
    int enqueued = 0;
    std::deque<hipEvent_t> inflight;
};

bool same_grid(const qvti_grid& a, const qvti_grid& b) { return memcmp(&a, &b, sizeof a) == 0; }

}  // namespace

struct qvti_server {
    int device = 0;
    long long budget = 0;
    int max_active = 4;
    int slice = 25;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> waiting;
    std::vector<std::shared_ptr<Request>> active;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    struct Pooled {
        qvti_grid grid;
        qvti_ctx* ctx;
        long long bytes;
    };
    std::vector<Pooled> pool;
    long long used = 0;  // active + pooled bytes
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void run();
    bool admit_one(std::unique_lock<std::mutex>& lk);
    void finish(const std::shared_ptr<Request>& r, int status);
};

static std::vector<float> copy_arr(const float* p, long long n)
{
    return p ? std::vector<float>(p, p + n) : std::vector<float>();
}

// Admit the head of the queue if a slot and memory allow; returns true if one
// was admitted (or rejected).  Called with the lock held; device work happens
// with the lock released.
bool qvti_server::admit_one(std::unique_lock<std::mutex>& lk)
{
    if (waiting.empty() || (int)active.size() >= max_active) return false;
    std::shared_ptr<Request> r = waiting.front();
    if (r->cancel) {
        waiting.pop_front();
        finish(r, QVTI_E_CANCELLED);
        return true;
    }
    const long long need = qvti_estimate_bytes(&r->grid, 1);
    if (need > budget) {
        waiting.pop_front();
        finish(r, QVTI_E_BUDGET);
        return true;
    }
    // a pooled context of the same grid?
    qvti_ctx* ctx = nullptr;
    long long bytes = 0;
    for (size_t i = 0; i < pool.size(); ++i)
        if (same_grid(pool[i].grid, r->grid)) {
            ctx = pool[i].ctx;
            bytes = pool[i].bytes;
            pool.erase(pool.begin() + i);
            break;
        }
    if (!ctx) {
        // evict pooled contexts of other shapes until the request fits
        while (used + need > budget && !pool.empty()) {
            Pooled p = pool.front();
            pool.erase(pool.begin());
            used -= p.bytes;
            lk.unlock();
            qvti_destroy(p.ctx);
            lk.lock();
        }
        if (used + need > budget) return false;  // wait for active requests to finish
        used += need;
        lk.unlock();
        int rc = qvti_create(&ctx, &r->grid, 1, device);
        lk.lock();
        if (rc) {
            used -= need;
            waiting.pop_front();
            finish(r, rc);
            return true;
        }
        bytes = need;
    }
    waiting.pop_front();
    r->ctx = ctx;
    r->bytes = bytes;
    r->t_admit = now();
    active.push_back(r);
    // upload the model and start the shot (device work, lock released)
    lk.unlock();
    const float* m[7];
    for (int k = 0; k < 7; ++k) m[k] = r->model[k].empty() ? nullptr : r->model[k].data();
    int rc = qvti_set_model(ctx, m[0], m[1], m[2], m[3], m[4], m[5], m[6]);
    if (!rc) {
        qvti_shot sh;
        sh.kind = r->kind;
        sh.sx = r->sx;
        sh.sy = r->sy;
        sh.sz = r->sz;
        sh.nrec = (int)r->rx.size();
        sh.rx = r->rx.data();
        sh.ry = r->ry.data();
        sh.rz = r->rz.data();
        rc = qvti_begin(ctx, r->nt, r->rik.data(), &sh, 1);
    }
    lk.lock();
    if (rc) finish(r, rc);
    return true;
}

// Retire an active (or never admitted) request.  Lock held.
void qvti_server::finish(const std::shared_ptr<Request>& r, int status)
{
    r->status = status;
    r->t_done = now();
    for (auto e : r->inflight) hipEventDestroy(e);
    r->inflight.clear();
    for (size_t i = 0; i < active.size(); ++i)
        if (active[i] == r) {
            active.erase(active.begin() + i);
            break;
        }
    if (r->ctx) {
        pool.push_back({r->grid, r->ctx, r->bytes});
        r->ctx = nullptr;
    }
    for (int k = 0; k < 7; ++k) std::vector<float>().swap(r->model[k]);
    cv.notify_all();
}

void qvti_server::run()
{
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (stop && waiting.empty() && active.empty()) break;
        while (admit_one(lk)) {
        }
        bool did = false;
        // retire finished slices, then top up every active request
        std::vector<std::shared_ptr<Request>> act = active;
        for (auto& r : act) {
            while (!r->inflight.empty() && hipEventQuery(r->inflight.front()) == hipSuccess) {
                hipEventDestroy(r->inflight.front());
                r->inflight.pop_front();
            }
            if (r->cancel) {
                lk.unlock();
                qvti_sync(r->ctx);
                lk.lock();
                finish(r, QVTI_E_CANCELLED);
                did = true;
                continue;
            }
            if (r->enqueued >= r->nt && r->inflight.empty()) {
                lk.unlock();
                int rc = qvti_finish(r->ctx, r->rec.data(), r->rec.data() + (size_t)r->nt * r->rx.size(),
                                     r->rec.data() + 2 * (size_t)r->nt * r->rx.size());
                lk.lock();
                finish(r, rc);
                did = true;
                continue;
            }
            if (r->enqueued < r->nt && r->inflight.size() < 2) {
                const int k = std::min(slice, r->nt - r->enqueued);
                lk.unlock();
                int rc = qvti_step(r->ctx, k);
                hipEvent_t ev = nullptr;
                if (!rc && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
                    // record on the context's stream: qvti_step enqueued there
                    hipEventRecord(ev, qvti_ctx_stream(r->ctx));
                }
                lk.lock();
                if (rc) {
                    finish(r, rc);
                } else {
                    r->enqueued += k;
                    if (ev) r->inflight.push_back(ev);
                }
                did = true;
            }
        }
        if (!did) cv.wait_for(lk, std::chrono::microseconds(200));
    }
    for (auto& p : pool) {
        lk.unlock();
        qvti_destroy(p.ctx);
        lk.lock();
    }
    pool.clear();
}

extern "C" int qvti_server_create(qvti_server** out, int device, long long hbm_budget_bytes, int max_active,
                                  int steps_per_slice)
{
    if (!out || hbm_budget_bytes <= 0 || max_active < 1 || steps_per_slice < 1)
        return qvti_fail(QVTI_E_ARG, "bad server arguments");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return qvti_fail(QVTI_E_DEVICE, "no HIP device %d", device);
    qvti_server* s = new qvti_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->slice = steps_per_slice;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return QVTI_OK;
}

extern "C" int qvti_server_submit(qvti_server* s, const qvti_request* q, uint64_t* id)
{
    if (!s || !q || !id) return qvti_fail(QVTI_E_ARG, "NULL argument");
    const qvti_grid& g = q->grid;
    const long long N = (long long)g.nx * g.ny * g.nz;
    if (N <= 0 || q->nt < 1 || !q->rik || !q->velp || !q->gama_p || !q->vels || !q->gama_s || !q->rho ||
        (g.vti && (!q->eps || !q->delta)) || q->shot.nrec < 0 ||
        (q->shot.nrec > 0 && (!q->shot.rx || !q->shot.ry || !q->shot.rz)))
        return qvti_fail(QVTI_E_ARG, "incomplete request");
    auto r = std::make_shared<Request>();
    r->grid = g;
    r->nt = q->nt;
    const float* m[7] = {q->velp, q->gama_p, q->vels, q->gama_s, q->rho, g.vti ? q->eps : nullptr,
                         g.vti ? q->delta : nullptr};
    for (int k = 0; k < 7; ++k) r->model[k] = copy_arr(m[k], m[k] ? N : 0);
    r->rik = copy_arr(q->rik, q->nt);
    r->kind = q->shot.kind;
    r->sx = q->shot.sx;
    r->sy = q->shot.sy;
    r->sz = q->shot.sz;
    r->rx.assign(q->shot.rx, q->shot.rx + q->shot.nrec);
    r->ry.assign(q->shot.ry, q->shot.ry + q->shot.nrec);
    r->rz.assign(q->shot.rz, q->shot.rz + q->shot.nrec);
    r->rec.assign((size_t)3 * q->nt * q->shot.nrec, 0.f);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return qvti_fail(QVTI_E_STATE, "server is shutting down");
    r->id = s->next_id++;
    r->t_submit = s->now();
    s->waiting.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    s->cv.notify_all();
    return QVTI_OK;
}

extern "C" int qvti_server_poll(qvti_server* s, uint64_t id, qvti_result_info* info, float* rec_vx, float* rec_vy,
                                float* rec_vz)
{
    if (!s) return qvti_fail(QVTI_E_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return qvti_fail(QVTI_E_ARG, "unknown request %llu", (unsigned long long)id);
    const Request& r = *it->second;
    if (info) {
        info->status = r.status;
        info->t_submit_s = r.t_submit;
        info->t_admit_s = r.t_admit;
        info->t_done_s = r.t_done;
        info->nt = r.nt;
        info->nrec = (int)r.rx.size();
    }
    if (r.status == QVTI_E_PENDING) return QVTI_E_PENDING;
    if (r.status == QVTI_OK) {
        const size_t n = (size_t)r.nt * r.rx.size();
        if (rec_vx) memcpy(rec_vx, r.rec.data(), n * sizeof(float));
        if (rec_vy) memcpy(rec_vy, r.rec.data() + n, n * sizeof(float));
        if (rec_vz) memcpy(rec_vz, r.rec.data() + 2 * n, n * sizeof(float));
    }
    return r.status == QVTI_OK ? QVTI_OK : qvti_fail(r.status, "request %llu ended with status %d",
                                                     (unsigned long long)id, r.status);
}

extern "C" int qvti_server_wait(qvti_server* s, uint64_t id, double timeout_s)
{
    if (!s) return qvti_fail(QVTI_E_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return qvti_fail(QVTI_E_ARG, "unknown request");
    auto r = it->second;
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (r->status == QVTI_E_PENDING)
        if (s->cv.wait_until(lk, until) == std::cv_status::timeout && r->status == QVTI_E_PENDING)
            return QVTI_E_PENDING;
    return r->status;
}

extern "C" int qvti_server_cancel(qvti_server* s, uint64_t id)
{
    if (!s) return qvti_fail(QVTI_E_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return qvti_fail(QVTI_E_ARG, "unknown request");
    it->second->cancel = true;
    s->cv.notify_all();
    return QVTI_OK;
}

extern "C" double qvti_server_clock(qvti_server* s) { return s ? s->now() : -1.0; }

extern "C" int qvti_server_destroy(qvti_server* s)
{
    if (!s) return QVTI_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto& r : s->waiting) r->cancel = true;
        for (auto& r : s->active) r->cancel = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    delete s;
    return QVTI_OK;
}
