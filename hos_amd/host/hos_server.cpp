// hos_server.cpp -- continuous-batching request loop (include/hos_serve.h).
// Host C++ on top of the native C ABI (include/hos_gpu.h).
#include "../../include/hos_serve.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/hos_gpu.h"

namespace {

using clk = std::chrono::steady_clock;

struct Req {
    hos_request r{};
    std::vector<double> u0;
    std::vector<double> result;
    int state = HOS_REQ_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double H = 0, vol = 0;
};

struct Pool {
    hos_request key{};
    hos_ctx *ctx = nullptr;
    int cap = 0, active = 0;
    int inflight = 0;        // steps enqueued by the last launch and not yet collected
    bool broken = false;     // a launch or its steps failed: no more admissions
    double empty_since = 0;  // server time the pool last became empty
    size_t bytes = 0;
    std::vector<long long> slot_req;
    std::vector<int> remaining;
};

bool same_shape(const hos_request &a, const hos_request &b)
{
    return a.Nx == b.Nx && a.Ny == b.Ny && a.nlevs == b.nlevs && a.Lx == b.Lx && a.Ly == b.Ly && a.g == b.g;
}

double pct(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (v.size() - 1);
    const size_t i = (size_t)std::floor(pos);
    const size_t j = std::min(i + 1, v.size() - 1);
    return v[i] + (pos - i) * (v[j] - v[i]);
}

}  // namespace

struct hos_server {
    hos_server_params p{};
    std::mutex mu;          // queue, requests, statistics
    std::mutex step_mu;     // one scheduler iteration at a time
    std::condition_variable cv;
    std::deque<long long> queue;
    std::map<long long, Req> reqs;
    std::vector<Pool *> pools;
    long long next_id = 1;
    size_t used = 0, peak = 0;
    std::thread worker;
    std::atomic<bool> stop{false}, started{false};
    clk::time_point t0 = clk::now();
    std::vector<double> step_ms, lat, qwait;
    double pt_steps = 0, busy = 0;
    int pools_created = 0, peak_active = 0, running = 0;
    double t_last = 0;        // time of the previous scheduler iteration
    int inject_fail = 0;      // pool launches still to fail (hos_server_fail_launches)
    bool any_inflight = false; // a pool had steps in flight after it

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

static hos_params pool_params(const hos_server *s, const hos_request &r)
{
    hos_params p;
    hos_params_default(&p);
    p.Nx = r.Nx;
    p.Ny = r.Ny;
    p.nlevs = r.nlevs;
    p.Lx = r.Lx;
    p.Ly = r.Ly;
    p.g = r.g;
    p.depth = 0.0;  // per member (hos_set_member)
    p.precision = s->p.precision;
    p.compensated = s->p.compensated;
    p.scheme = 2;
    return p;
}

static int default_cap(const hos_server *s, const hos_request &r)
{
    if (s->p.max_slots > 0) return s->p.max_slots;
    const long pts = (long)r.Nx * r.Ny;
    return (int)std::max(2L, std::min(64L, (1L << 23) / pts));
}

extern "C" int hos_server_create(const hos_server_params *p, hos_server **out)
{
    if (!p || !out) return HOS_E_INVALID;
    if (!(p->hbm_budget > 0) || (p->precision != HOS_FP32 && p->precision != HOS_FP64) || p->max_slots < 0 ||
        p->steps_per_launch < 0)
        return HOS_E_INVALID;
    hos_server *s = new hos_server();
    s->p = *p;
    if (s->p.steps_per_launch == 0) s->p.steps_per_launch = 8;
    *out = s;
    return HOS_OK;
}

extern "C" int hos_server_submit(hos_server *s, const hos_request *r, const double *u0, long long *id)
{
    if (!s || !r || !u0 || !id) return HOS_E_INVALID;
    hos_params pp = pool_params(s, *r);
    const size_t one = hos_footprint(&pp, 1);
    if (one == 0 || r->nsteps < 0 || !(r->dt > 0) || one > (size_t)s->p.hbm_budget) return HOS_E_INVALID;
    Req q;
    q.r = *r;
    const size_t n = (size_t)4 * r->Nx * (r->Ny / 2 + 1);
    q.u0.assign(u0, u0 + n);
    std::lock_guard<std::mutex> g(s->mu);
    q.t_submit = s->now();
    *id = s->next_id++;
    if (r->nsteps == 0) {  // nothing to run: the result is the initial state
        q.result = q.u0;
        q.state = HOS_REQ_DONE;
        q.t_admit = q.t_done = q.t_submit;
        s->reqs.emplace(*id, std::move(q));
        return HOS_OK;
    }
    s->reqs.emplace(*id, std::move(q));
    s->queue.push_back(*id);
    s->cv.notify_all();
    return HOS_OK;
}

extern "C" int hos_server_status(hos_server *s, long long id)
{
    if (!s) return HOS_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? HOS_E_INVALID : it->second.state;
}

extern "C" int hos_server_result(hos_server *s, long long id, double *u, double *info)
{
    if (!s) return HOS_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return HOS_E_INVALID;
    Req &q = it->second;
    if (q.state != HOS_REQ_DONE && q.state != HOS_REQ_FAILED) return HOS_E_STATE;
    if (u && q.state == HOS_REQ_DONE) std::memcpy(u, q.result.data(), q.result.size() * sizeof(double));
    if (info) {
        const double rec[6] = {q.t_submit, q.t_admit, q.t_done, (double)q.r.nsteps, q.H, q.vol};
        std::memcpy(info, rec, sizeof rec);
    }
    const int st = q.state;
    s->reqs.erase(it);
    return st == HOS_REQ_DONE ? HOS_OK : HOS_E_DEVICE;
}

static void destroy_pool(hos_server *s, size_t i)
{
    Pool *p = s->pools[i];
    hos_destroy(p->ctx);
    s->used -= p->bytes;
    delete p;
    s->pools.erase(s->pools.begin() + (long)i);
}

// release empty pools: all of them, or those empty for longer than a second
static void release_idle(hos_server *s, bool all)
{
    const double now = s->now();
    for (size_t i = 0; i < s->pools.size();) {
        Pool *p = s->pools[i];
        if (p->active == 0 && p->inflight == 0 && (all || p->broken || now - p->empty_since > 1.0))
            destroy_pool(s, i);
        else
            ++i;
    }
}

// find an idle pool with a free slot for r, creating one if the budget allows
// (after releasing empty pools, if it does not at first); nullptr: wait
static Pool *pool_for(hos_server *s, const hos_request &r)
{
    bool same_busy = false, same_full = false;
    for (Pool *p : s->pools) {
        if (!same_shape(p->key, r) || p->broken) continue;
        if (p->inflight > 0) {
            same_busy |= p->active < p->cap;
            continue;
        }
        if (p->active < p->cap) return p;
        same_full = true;
    }
    if (same_busy || same_full) return nullptr;  // a slot of this shape frees up soon
    hos_params pp = pool_params(s, r);
    for (int pass = 0; pass < 2; ++pass) {
        for (int cap = default_cap(s, r); cap >= 1; cap /= 2) {
            const size_t b = hos_footprint(&pp, cap);
            if (b == 0 || s->used + b > (size_t)s->p.hbm_budget) continue;
            hos_ctx *c = nullptr;
            if (hos_create(&pp, cap, &c) != HOS_OK) return nullptr;
            hos_use_graph(c, s->p.use_graph);
            hos_set_active(c, 0);
            Pool *p = new Pool();
            p->key = r;
            p->ctx = c;
            p->cap = cap;
            p->bytes = b;
            p->slot_req.assign(cap, 0);
            p->remaining.assign(cap, 0);
            s->pools.push_back(p);
            s->used += b;
            s->peak = std::max(s->peak, s->used);
            s->pools_created++;
            return p;
        }
        release_idle(s, true);  // empty pools of other shapes give their memory back
    }
    return nullptr;
}

// mark the requests of a pool's active members failed (none of their steps
// in flight counts) and empty the pool
static void fail_pool(hos_server *s, Pool *p)
{
    std::lock_guard<std::mutex> g(s->mu);
    for (int k = 0; k < p->active; ++k) {
        Req &q = s->reqs[p->slot_req[k]];
        q.state = HOS_REQ_FAILED;
        q.t_done = s->now();
        s->lat.push_back(q.t_done - q.t_submit);
        s->running--;
    }
    p->active = 0;
    p->inflight = 0;
    p->broken = true;
    p->empty_since = s->now();
}

// retire the finished members of an idle pool, refilling their slots from the
// pool's end; returns HOS_OK or the first error
static int retire(hos_server *s, Pool *p)
{
    int rc = HOS_OK;
    for (int k = p->active - 1; k >= 0; --k) {
        if (p->remaining[k] > 0) continue;
        const long long id = p->slot_req[k];
        std::vector<double> u((size_t)4 * p->key.Nx * (p->key.Ny / 2 + 1));
        double H = 0, vol = 0;
        int e1 = hos_get_state(p->ctx, k, u.data());
        if (e1 == HOS_OK) e1 = hos_get_energy(p->ctx, k, &H, &vol);
        {
            std::lock_guard<std::mutex> g(s->mu);
            Req &q = s->reqs[id];
            q.result.swap(u);
            q.H = H;
            q.vol = vol;
            q.state = e1 == HOS_OK ? HOS_REQ_DONE : HOS_REQ_FAILED;
            q.t_done = s->now();
            s->lat.push_back(q.t_done - q.t_submit);
            s->running--;
        }
        const int last = p->active - 1;
        if (k != last) {
            if (hos_copy_member(p->ctx, last, k) != HOS_OK) rc = HOS_E_DEVICE;
            p->slot_req[k] = p->slot_req[last];
            p->remaining[k] = p->remaining[last];
        }
        p->active--;
        hos_set_active(p->ctx, p->active);
        if (p->active == 0) p->empty_since = s->now();
    }
    return rc;
}

extern "C" int hos_server_step(hos_server *s, int *active)
{
    if (!s) return HOS_E_INVALID;
    std::lock_guard<std::mutex> sg(s->step_mu);
    int rc = HOS_OK;
    {
        // wall time with steps on the device (between this iteration and the last)
        const double tn = s->now();
        std::lock_guard<std::mutex> g(s->mu);
        if (s->any_inflight) s->busy += tn - s->t_last;
        s->t_last = tn;
    }
    // 1. collect the pools whose steps have completed
    for (size_t i = 0; i < s->pools.size(); ++i) {
        Pool *p = s->pools[i];
        if (p->inflight == 0) continue;
        const int q = hos_query(p->ctx);
        if (q == 1) continue;
        if (q < 0) {
            rc = HOS_E_DEVICE;
            fail_pool(s, p);
            continue;
        }
        const int n = hos_step_times(p->ctx, nullptr, 0);
        std::vector<double> ms(n > 0 ? n : 0);
        if (n > 0) hos_step_times(p->ctx, ms.data(), n);
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->step_ms.insert(s->step_ms.end(), ms.begin(), ms.end());
            s->pt_steps += (double)p->active * p->key.Nx * p->key.Ny * p->inflight;
        }
        for (int k = 0; k < p->active; ++k) p->remaining[k] -= p->inflight;
        p->inflight = 0;
    }
    // 2. retire finished members of the idle pools
    for (Pool *p : s->pools)
        if (p->inflight == 0 && p->active > 0 && retire(s, p) != HOS_OK) rc = HOS_E_DEVICE;
    // 3. admission into free slots of idle pools
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            Req &q = s->reqs[*it];
            Pool *p = pool_for(s, q.r);
            if (!p) {
                if (s->pools.empty()) {  // nothing will ever be released: cannot fit
                    q.state = HOS_REQ_FAILED;
                    q.t_done = s->now();
                    it = s->queue.erase(it);
                    continue;
                }
                ++it;
                continue;
            }
            const int slot = p->active;
            if (hos_set_state(p->ctx, slot, q.u0.data()) != HOS_OK ||
                hos_set_member(p->ctx, slot, q.r.depth, q.r.dt, 0.0) != HOS_OK) {
                q.state = HOS_REQ_FAILED;
                q.t_done = s->now();
                it = s->queue.erase(it);
                continue;
            }
            p->slot_req[slot] = *it;
            p->remaining[slot] = q.r.nsteps;
            p->active++;
            hos_set_active(p->ctx, p->active);
            q.state = HOS_REQ_RUNNING;
            q.t_admit = s->now();
            s->qwait.push_back(q.t_admit - q.t_submit);
            s->running++;
            it = s->queue.erase(it);
        }
    }
    // 4. launch every idle pool with active members: k steps in flight on its
    //    own stream, k = the fewest steps any member still needs (capped)
    int nact = 0;
    bool inflight = false;
    for (Pool *p : s->pools) {
        if (p->inflight == 0 && p->active > 0) {
            int k = s->p.steps_per_launch;
            for (int m = 0; m < p->active; ++m) k = std::min(k, p->remaining[m]);
            if (k < 1) k = 1;  // (retire leaves no member with 0 steps to go)
            const bool injected = s->inject_fail > 0 && s->inject_fail-- > 0;
            if (injected || hos_step_async(p->ctx, k) != HOS_OK) {
                rc = HOS_E_DEVICE;
                fail_pool(s, p);
            } else {
                p->inflight = k;
            }
        }
        nact += p->active;
        inflight |= p->inflight > 0;
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        release_idle(s, false);
        s->peak_active = std::max(s->peak_active, nact);
        s->any_inflight = inflight;
    }
    if (active) *active = nact;
    s->cv.notify_all();
    return rc;
}

extern "C" int hos_server_start(hos_server *s)
{
    if (!s) return HOS_E_INVALID;
    if (s->started.exchange(true)) return HOS_E_STATE;
    s->stop = false;
    s->worker = std::thread([s] {
        while (!s->stop) {
            bool work;
            {
                std::unique_lock<std::mutex> g(s->mu);
                work = !s->queue.empty() || s->running > 0;
                if (!work) s->cv.wait_for(g, std::chrono::milliseconds(5));
                work = !s->queue.empty() || s->running > 0;
            }
            if (work) {
                hos_server_step(s, nullptr);
                // pools with steps in flight are polled, not waited for
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
        }
    });
    return HOS_OK;
}

extern "C" int hos_server_stop(hos_server *s)
{
    if (!s) return HOS_E_INVALID;
    if (!s->started) return HOS_OK;
    s->stop = true;
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    s->started = false;
    return HOS_OK;
}

extern "C" int hos_server_drain(hos_server *s, double timeout_s)
{
    if (!s) return HOS_E_INVALID;
    const auto end = clk::now() + std::chrono::duration<double>(timeout_s);
    for (;;) {
        {
            std::unique_lock<std::mutex> g(s->mu);
            if (s->queue.empty() && s->running == 0) return HOS_OK;
            if (clk::now() >= end) return HOS_E_STATE;
            if (s->started) {
                s->cv.wait_for(g, std::chrono::milliseconds(2));
                continue;
            }
        }
        hos_server_step(s, nullptr);
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

extern "C" int hos_server_stats(hos_server *s, double *out, int n)
{
    if (!s || !out) return HOS_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    const double v[12] = {pct(s->step_ms, 50),
                          pct(s->step_ms, 99),
                          pct(s->lat, 50),
                          pct(s->lat, 99),
                          (double)s->lat.size(),
                          s->pt_steps,
                          s->busy,
                          (double)s->pools_created,
                          (double)s->peak,
                          (double)s->peak_active,
                          pct(s->qwait, 50),
                          pct(s->qwait, 99)};
    for (int i = 0; i < n && i < 12; ++i) out[i] = v[i];
    return HOS_OK;
}

extern "C" int hos_server_fail_launches(hos_server *s, int n)
{
    if (!s || n < 0) return HOS_E_INVALID;
    std::lock_guard<std::mutex> sg(s->step_mu);
    s->inject_fail = n;
    return HOS_OK;
}

extern "C" int hos_server_destroy(hos_server *s)
{
    if (!s) return HOS_OK;
    hos_server_stop(s);
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (Pool *p : s->pools) {
            hos_destroy(p->ctx);
            delete p;
        }
        s->pools.clear();
    }
    delete s;
    return HOS_OK;
}
