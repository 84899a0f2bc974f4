// Continuous-batching HOLD server (host runtime, no kernels of its own).
//
// No reference counterpart.  Requests (one Plummer-type system each: host
// arrays, t_end, dt_root) are admitted in arrival order while their device
// footprint fits the HBM budget; admitted work starts as soon as an
// execution slot is free, so requests that arrive while others run join the
// device as slots and memory free up:
//  * small systems (n <= small_max_n) are gathered into a batch and run by
//    one persistent ensemble launch (hhs_ensemble_run_f64) on one of
//    `batch_streams` streams -- a new batch can start while older ones run,
//    and each request of a batch is handed back as soon as its own system
//    is done (per-system flags in host-mapped memory);
//  * large systems run the per-system HOLD driver (hhs_hold_f64) on one of
//    `big_workers` worker threads, each with its own stream and a reusable
//    hhs_globals sized for max_n.
// All streams are non-blocking and at the default priority; each batch
// stream keeps grow-only device (stream-ordered) and host-mapped buffers,
// so batches allocate nothing once the buffers have grown and nothing is
// allocated or freed while the scheduler's lock is held; no call forces a
// device-wide synchronisation while other streams are busy.  A request's
// result depends only on its own data (the path is chosen by n), so it is
// the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hhs_internal.h"

namespace hhs {
namespace {

enum ReqState { kPending = 0, kAdmitted = 1, kRunning = 2, kDone = 3, kFailed = 4 };

struct Request {
    long long ticket = 0;
    int n = 0;
    double t_end = 0.0, dt_root = 0.0;
    std::vector<double> pos, vel, m;
    size_t footprint = 0;
    bool small = true;
    hhs_request_info info{};
};

// One batch stream with its buffers.  The buffers only grow and are
// reused by every batch of the slot: the device block (inputs, workspace)
// through the stream-ordered allocator on the slot's stream, the
// host-mapped block (pos/vel in and out, results, flags) geometrically,
// keeping outgrown host blocks until the server is destroyed -- so no batch
// frees host memory, whose free would synchronise the whole device.  Only
// launch_batch (no lock held) allocates; a slot's next batch starts after
// the previous one has completed and been handed back.
struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    void *dev = nullptr;
    size_t dev_bytes = 0;
    void *host = nullptr, *host_dev = nullptr;  // host address / device address
    size_t host_bytes = 0;
    std::vector<void *> retired_host;
};

struct Batch {
    std::vector<Request *> reqs;
    std::vector<size_t> first;    // particle offset of each request
    std::vector<char> finished;   // request already handed back
    int slot = -1;
    size_t n_part = 0;
    double *h_pos = nullptr, *h_vel = nullptr;
    hhs_system_result *h_res = nullptr;
    int *h_done = nullptr;
};

size_t small_footprint(int n) {
    // packed pos/vel/m + ensemble working and scratch copies
    return static_cast<size_t>(n) * (56 + 136) + 512;
}
size_t big_footprint(int n) {
    // particle set (96 B) + slow/fast level sets (192 B per level), 24 levels
    return static_cast<size_t>(n) * (96 + 192 * 24) + 4096;
}

}  // namespace
}  // namespace hhs

struct hhs_server {
    hhs_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<hhs::Request *> pending;
    std::unordered_map<long long, std::unique_ptr<hhs::Request>> all;
    std::vector<hhs::Request *> admitted_small;
    std::deque<hhs::Request *> admitted_big;
    std::vector<hhs::Slot> slots;
    std::vector<bool> slot_busy;
    std::vector<hhs::Batch *> inflight;
    std::vector<std::thread> workers;
    std::thread scheduler;
    size_t used = 0, peak = 0;
    long long next_ticket = 1;
    int big_running = 0;
    bool stop = false;          // no new work after the queues drain
    bool workers_stop = false;  // set once the scheduler has exited
    int device = 0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace hhs {
namespace {

void fail_request(hhs_server *s, Request *r, int status) {
    r->info.result.status = status;
    r->info.state = kFailed;
    r->info.t_done = s->now();
    s->used -= r->footprint;
}

// Launch one ensemble batch (called without the lock held).  Positions,
// velocities, results and per-system done flags live in host-mapped
// coherent memory: the kernel reads the initial state from it, writes each
// system's final state and result there and then raises that system's flag,
// so the scheduler can hand back every request as soon as its own system
// is done instead of when the whole batch is.
int grow_slot(Slot &sl, size_t dev_bytes, size_t host_bytes) {
    if (dev_bytes > sl.dev_bytes) {
        const size_t want = std::max(dev_bytes, 2 * sl.dev_bytes);
        if (sl.dev) HHS_HIP(hipFreeAsync(sl.dev, sl.st));
        sl.dev = nullptr;
        sl.dev_bytes = 0;
        HHS_HIP(hipMallocAsync(&sl.dev, want, sl.st));
        sl.dev_bytes = want;
    }
    if (host_bytes > sl.host_bytes) {
        const size_t want = std::max(host_bytes, 2 * sl.host_bytes);
        if (sl.host) sl.retired_host.push_back(sl.host);
        sl.host = sl.host_dev = nullptr;
        sl.host_bytes = 0;
        HHS_HIP(hipHostMalloc(&sl.host, want, hipHostMallocMapped | hipHostMallocCoherent));
        sl.host_bytes = want;
        HHS_HIP(hipHostGetDevicePointer(&sl.host_dev, sl.host, 0));
    }
    return HHS_OK;
}

int launch_batch(hhs_server *s, Batch *b) {
    const int k = static_cast<int>(b->reqs.size());
    Slot &sl = s->slots[b->slot];
    hipStream_t st = sl.st;
    std::vector<int> off(k + 1, 0);
    for (int q = 0; q < k; ++q) off[q + 1] = off[q] + b->reqs[q]->n;
    const size_t np = static_cast<size_t>(off[k]);
    b->n_part = np;
    b->first.assign(off.begin(), off.end() - 1);
    b->finished.assign(k, 0);
    const size_t wsb = hhs_ensemble_workspace_bytes(k, static_cast<long long>(np));
    auto al = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    // device: offsets, m, t_end, dt_root, workspace
    const size_t o_off = 0, o_m = o_off + al((k + 1) * 4), o_te = o_m + al(np * 8 + 8),
                 o_dt = o_te + al(k * 8), o_ws = o_dt + al(k * 8), dev_bytes = o_ws + al(wsb);
    // host-mapped: staging of the device inputs, then pos, vel, results, flags
    const size_t h_pos = o_ws, h_vel = h_pos + al(np * 24 + 8), h_res = h_vel + al(np * 24 + 8),
                 h_done = h_res + al(k * sizeof(hhs_system_result)),
                 h_bytes = h_done + al(k * sizeof(int));
    if (int rc = grow_slot(sl, dev_bytes, h_bytes)) return rc;
    char *h = static_cast<char *>(sl.host);
    std::memcpy(h + o_off, off.data(), (k + 1) * 4);
    double *hp = reinterpret_cast<double *>(h + h_pos), *hv = reinterpret_cast<double *>(h + h_vel),
           *hm = reinterpret_cast<double *>(h + o_m), *hte = reinterpret_cast<double *>(h + o_te),
           *hdt = reinterpret_cast<double *>(h + o_dt);
    for (int q = 0; q < k; ++q) {
        const Request *r = b->reqs[q];
        std::memcpy(hp + 3 * off[q], r->pos.data(), r->n * 24);
        std::memcpy(hv + 3 * off[q], r->vel.data(), r->n * 24);
        std::memcpy(hm + off[q], r->m.data(), r->n * 8);
        hte[q] = r->t_end;
        hdt[q] = r->dt_root;
    }
    b->h_pos = hp;
    b->h_vel = hv;
    b->h_res = reinterpret_cast<hhs_system_result *>(h + h_res);
    b->h_done = reinterpret_cast<int *>(h + h_done);
    std::memset(b->h_done, 0, k * sizeof(int));
    char *hd = static_cast<char *>(sl.host_dev);
    char *d = static_cast<char *>(sl.dev);
    HHS_HIP(hipMemcpyAsync(d, h, o_ws, hipMemcpyHostToDevice, st));
    hhs_ensemble_desc desc{};
    desc.n_systems = k;
    desc.offsets = reinterpret_cast<const int *>(d + o_off);
    desc.pos = reinterpret_cast<double *>(hd + h_pos);
    desc.vel = reinterpret_cast<double *>(hd + h_vel);
    desc.m = reinterpret_cast<const double *>(d + o_m);
    desc.t_end = reinterpret_cast<const double *>(d + o_te);
    desc.dt_root = reinterpret_cast<const double *>(d + o_dt);
    desc.eta = s->cfg.eta;
    desc.eps = s->cfg.eps;
    desc.max_level = s->cfg.max_level;
    desc.max_calls = 0;
    desc.wave_team_max_n = 0;
    desc.results = reinterpret_cast<hhs_system_result *>(hd + h_res);
    desc.timestep_mode = s->cfg.timestep_mode;
    desc.done = reinterpret_cast<int *>(hd + h_done);
    int rc = hhs_ensemble_run_f64(&desc, off.data(), d + o_ws, wsb, st);
    if (rc) return rc;
    HHS_HIP(hipEventRecord(sl.done, st));
    return HHS_OK;
}

// Hand back request q of a batch whose flag is up (lock held).
void finish_request(hhs_server *s, Batch *b, size_t q, double t) {
    Request *r = b->reqs[q];
    std::memcpy(r->pos.data(), b->h_pos + 3 * b->first[q], r->n * 24);
    std::memcpy(r->vel.data(), b->h_vel + 3 * b->first[q], r->n * 24);
    r->info.result = b->h_res[q];
    r->info.state = b->h_res[q].status == HHS_OK ? kDone : kFailed;
    r->info.t_done = t;
    s->used -= r->footprint;
    b->finished[q] = 1;
}

// Requests of a running batch whose systems are done (lock held).
bool poll_batch(hhs_server *s, Batch *b) {
    bool any = false;
    const double t = s->now();
    for (size_t q = 0; q < b->reqs.size(); ++q) {
        if (b->finished[q]) continue;
        if (__atomic_load_n(&b->h_done[q], __ATOMIC_ACQUIRE) == 1) {
            finish_request(s, b, q, t);
            any = true;
        }
    }
    return any;
}

// Retire a batch after its launch has completed (lock held; no HIP call).
// A failed launch_batch may have queued work on the slot's stream before it
// failed; the scheduler waits for the slot's stream before reusing it.
void finish_batch(hhs_server *s, Batch *b, int rc) {
    const double t = s->now();
    for (size_t q = 0; q < b->reqs.size(); ++q) {
        if (q < b->finished.size() && b->finished[q]) continue;
        if (rc == HHS_OK && b->h_done && __atomic_load_n(&b->h_done[q], __ATOMIC_ACQUIRE) == 1)
            finish_request(s, b, q, t);
        else
            fail_request(s, b->reqs[q], rc == HHS_OK ? HHS_ERR_HIP : rc);
    }
    s->slot_busy[b->slot] = false;
    delete b;
}

void scheduler_loop(hhs_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // 1. admission, strictly in arrival order, within the HBM budget
        while (!s->pending.empty()) {
            Request *r = s->pending.front();
            if (s->used + r->footprint > s->cfg.hbm_budget_bytes) break;
            s->pending.pop_front();
            s->used += r->footprint;
            s->peak = std::max(s->peak, s->used);
            r->info.state = kAdmitted;
            r->info.t_admit = s->now();
            if (r->small)
                s->admitted_small.push_back(r);
            else
                s->admitted_big.push_back(r);
            s->cv.notify_all();
        }
        // 2. start batches on free batch streams
        for (size_t slot = 0; slot < s->slot_busy.size() && !s->admitted_small.empty(); ++slot) {
            if (s->slot_busy[slot]) continue;
            Batch *b = new Batch();
            b->slot = static_cast<int>(slot);
            const size_t take =
                std::min<size_t>(s->admitted_small.size(), static_cast<size_t>(s->cfg.max_batch));
            b->reqs.assign(s->admitted_small.begin(), s->admitted_small.begin() + take);
            s->admitted_small.erase(s->admitted_small.begin(), s->admitted_small.begin() + take);
            const double t = s->now();
            for (Request *r : b->reqs) {
                r->info.state = kRunning;
                r->info.t_start = t;
            }
            s->slot_busy[slot] = true;
            lk.unlock();
            const int rc = launch_batch(s, b);
            // a launch that failed part-way leaves its queued work behind:
            // let it drain before the slot's buffers are reused
            if (rc) (void)hipStreamSynchronize(s->slots[slot].st);
            lk.lock();
            if (rc)
                finish_batch(s, b, rc);
            else
                s->inflight.push_back(b);
        }
        // 3. completions: single systems as their flags come up, then whole
        //    launches
        for (size_t q = 0; q < s->inflight.size();) {
            Batch *b = s->inflight[q];
            if (poll_batch(s, b)) s->cv.notify_all();
            const hipError_t e = hipEventQuery(s->slots[b->slot].done);
            if (e == hipErrorNotReady) {
                ++q;
                continue;
            }
            s->inflight.erase(s->inflight.begin() + q);
            finish_batch(s, b, e == hipSuccess ? HHS_OK : HHS_ERR_HIP);
            s->cv.notify_all();
        }
        const bool idle = s->pending.empty() && s->admitted_small.empty() &&
                          s->admitted_big.empty() && s->inflight.empty() && s->big_running == 0;
        if (s->stop && idle) break;
        s->cv.wait_for(lk, std::chrono::microseconds(200));
    }
}

void worker_loop(hhs_server *s) {
    (void)hipSetDevice(s->device);
    hipStream_t st = nullptr;
    hhs_globals *g = nullptr;
    void *dev = nullptr;
    const int nmax = s->cfg.max_n;
    const size_t nn = static_cast<size_t>(nmax);
    bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
              hhs_globals_create(&g, nmax, s->cfg.eta, s->cfg.eps, s->cfg.max_level) == HHS_OK &&
              hipMalloc(&dev, nn * 96 + 1024) == hipSuccess;
    if (ok) ok = hhs_globals_set_timestep_mode(g, s->cfg.timestep_mode) == HHS_OK;
    double *dpos = static_cast<double *>(dev);
    double *dvel = dpos + 3 * nn, *dacc = dvel + 3 * nn, *dts = dacc + 3 * nn, *dm = dts + nn;
    int *did = reinterpret_cast<int *>(dm + nn), *dpar = did + nn;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return s->workers_stop || !s->admitted_big.empty(); });
        if (s->admitted_big.empty()) break;  // workers_stop and nothing left
        Request *r = s->admitted_big.front();
        s->admitted_big.pop_front();
        r->info.state = kRunning;
        r->info.t_start = s->now();
        s->big_running++;
        lk.unlock();
        int rc = ok ? HHS_OK : HHS_ERR_ALLOC;
        hhs_system_result res{};
        const size_t n = static_cast<size_t>(r->n);
        if (rc == HHS_OK) {
            hhs_particles_f64 p{dpos, dvel, dacc, dts, dm, did, dpar, r->n};
            if (hipMemcpyAsync(dpos, r->pos.data(), n * 24, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(dvel, r->vel.data(), n * 24, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(dm, r->m.data(), n * 8, hipMemcpyHostToDevice, st))
                rc = HHS_ERR_HIP;
            double k0 = 0, u0 = 0, k1 = 0, u1 = 0;
            const double eps2 = s->cfg.eps * s->cfg.eps;
            if (!rc) rc = hhs_energy_f64(dpos, dvel, dm, r->n, eps2, &k0, &u0, st);
            if (!rc) {
                hhs_globals_reset_stats(g);
                rc = hhs_hold_f64(g, &p, r->t_end, r->dt_root, st);
            }
            if (!rc) rc = hhs_energy_f64(dpos, dvel, dm, r->n, eps2, &k1, &u1, st);
            if (!rc && (hipMemcpyAsync(r->pos.data(), dpos, n * 24, hipMemcpyDeviceToHost, st) ||
                        hipMemcpyAsync(r->vel.data(), dvel, n * 24, hipMemcpyDeviceToHost, st) ||
                        hipStreamSynchronize(st)))
                rc = HHS_ERR_HIP;
            hhs_hold_stats hs{};
            hhs_globals_stats(g, &hs);
            res.e0 = k0 + u0;
            res.e1 = k1 + u1;
            double t_reached = 0.0;
            while (t_reached < r->t_end) t_reached += r->dt_root;
            res.t = t_reached;
            res.calls = hs.calls;
            res.interactions = hs.interactions;
            res.ts_pairs = hs.ts_pairs;
            res.max_level = hs.max_level;
            res.status = rc;
        }
        lk.lock();
        s->big_running--;
        if (rc == HHS_OK) {
            r->info.result = res;
            r->info.state = kDone;
            r->info.t_done = s->now();
            s->used -= r->footprint;
        } else {
            fail_request(s, r, rc);
        }
        s->cv.notify_all();
    }
    lk.unlock();
    if (dev) (void)hipFree(dev);
    if (g) hhs_globals_destroy(g);
    if (st) (void)hipStreamDestroy(st);
}

}  // namespace
}  // namespace hhs

using namespace hhs;

extern "C" {

int hhs_server_create(hhs_server **out, const hhs_server_config *cfg) {
    HHS_REQUIRE(out && cfg, "server_create: null argument");
    *out = nullptr;
    HHS_REQUIRE(cfg->eta > 0.0 && cfg->eps >= 0.0, "server_create: need eta > 0, eps >= 0");
    HHS_REQUIRE(cfg->hbm_budget_bytes > 0, "server_create: zero HBM budget");
    HHS_REQUIRE(cfg->timestep_mode == HHS_TIMESTEP_EXACT || cfg->timestep_mode == HHS_TIMESTEP_FAST ||
                    cfg->timestep_mode == HHS_TIMESTEP_FP32,
                "server_create: bad timestep_mode");
    auto *s = new (std::nothrow) hhs_server();
    if (!s) return HHS_ERR_ALLOC;
    s->cfg = *cfg;
    // Defaults from the MI355X sweep of a Poisson stream (N log-uniform in
    // [64, 4096], t_end in [0.1, 1], lambda = 4 and 16 per s): N <= 512 in
    // ensemble batches on 2 streams and 4 per-system workers gave the lowest
    // p99 (1.9 s / 3.2 s) -- larger systems finish sooner on the whole GPU
    // (with device subtrees) than on one workgroup team of a batch.
    if (s->cfg.small_max_n <= 0) s->cfg.small_max_n = 512;
    if (s->cfg.max_n <= 0) s->cfg.max_n = 65536;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 4096;
    if (s->cfg.big_workers <= 0) s->cfg.big_workers = 4;
    if (s->cfg.batch_streams <= 0) s->cfg.batch_streams = 2;
    if (s->cfg.max_level <= 0) s->cfg.max_level = 40;
    s->t0 = std::chrono::steady_clock::now();
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    s->slots.resize(s->cfg.batch_streams);
    s->slot_busy.assign(s->cfg.batch_streams, false);
    for (auto &sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            int rc = hip_status(e, "server_create: stream", __FILE__, __LINE__);
            for (auto &x : s->slots) {
                if (x.done) (void)hipEventDestroy(x.done);
                if (x.st) (void)hipStreamDestroy(x.st);
            }
            delete s;
            return rc;
        }
    }
    s->scheduler = std::thread(scheduler_loop, s);
    for (int w = 0; w < s->cfg.big_workers; ++w) s->workers.emplace_back(worker_loop, s);
    *out = s;
    return HHS_OK;
}

int hhs_server_destroy(hhs_server *s) {
    if (!s) return HHS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->scheduler.join();  // returns once every admitted request has finished
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->workers_stop = true;
    }
    s->cv.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &sl : s->slots) {
        if (sl.dev) (void)hipFreeAsync(sl.dev, sl.st);
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.host) (void)hipHostFree(sl.host);
        for (void *h : sl.retired_host) (void)hipHostFree(h);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    delete s;
    return HHS_OK;
}

int hhs_server_submit(hhs_server *s, int n, const double *pos, const double *vel, const double *m,
                      double t_end, double dt_root, long long *ticket) {
    HHS_REQUIRE(s && ticket, "server_submit: null argument");
    HHS_REQUIRE(n > 0 && pos && vel && m, "server_submit: need n > 0 and host arrays");
    HHS_REQUIRE(dt_root > 0.0 && t_end >= 0.0, "server_submit: need dt_root > 0, t_end >= 0");
    auto r = std::make_unique<Request>();
    r->n = n;
    r->t_end = t_end;
    r->dt_root = dt_root;
    r->pos.assign(pos, pos + 3 * static_cast<size_t>(n));
    r->vel.assign(vel, vel + 3 * static_cast<size_t>(n));
    r->m.assign(m, m + n);
    std::lock_guard<std::mutex> lk(s->mu);
    r->small = n <= s->cfg.small_max_n;
    HHS_REQUIRE(r->small || n <= s->cfg.max_n, "server_submit: n=%d above max_n=%d", n,
                s->cfg.max_n);
    r->footprint = r->small ? small_footprint(n) : big_footprint(n);
    if (r->footprint > s->cfg.hbm_budget_bytes) {
        set_error("server_submit: request needs %zu bytes, budget is %zu", r->footprint,
                  s->cfg.hbm_budget_bytes);
        return HHS_ERR_BUDGET;
    }
    r->ticket = s->next_ticket++;
    r->info.state = kPending;
    r->info.path = r->small ? 0 : 1;
    r->info.t_submit = s->now();
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->all.emplace(r->ticket, std::move(r));
    s->cv.notify_all();
    return HHS_OK;
}

int hhs_server_poll(hhs_server *s, long long ticket, hhs_request_info *info) {
    HHS_REQUIRE(s && info, "server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    HHS_REQUIRE(it != s->all.end(), "server_poll: unknown ticket %lld", ticket);
    *info = it->second->info;
    return HHS_OK;
}

int hhs_server_wait(hhs_server *s, long long ticket, hhs_request_info *info) {
    HHS_REQUIRE(s, "server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    HHS_REQUIRE(it != s->all.end(), "server_wait: unknown ticket %lld", ticket);
    Request *r = it->second.get();
    s->cv.wait(lk, [&] { return r->info.state == kDone || r->info.state == kFailed; });
    if (info) *info = r->info;
    return r->info.state == kDone ? HHS_OK : r->info.result.status;
}

int hhs_server_fetch(hhs_server *s, long long ticket, double *pos_out, double *vel_out) {
    HHS_REQUIRE(s && pos_out && vel_out, "server_fetch: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    HHS_REQUIRE(it != s->all.end(), "server_fetch: unknown ticket %lld", ticket);
    Request *r = it->second.get();
    HHS_REQUIRE(r->info.state == kDone, "server_fetch: request %lld not done", ticket);
    std::memcpy(pos_out, r->pos.data(), r->pos.size() * 8);
    std::memcpy(vel_out, r->vel.data(), r->vel.size() * 8);
    return HHS_OK;
}

int hhs_server_release(hhs_server *s, long long ticket) {
    HHS_REQUIRE(s, "server_release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    HHS_REQUIRE(it != s->all.end(), "server_release: unknown ticket %lld", ticket);
    const int st = it->second->info.state;
    HHS_REQUIRE(st == kDone || st == kFailed, "server_release: request %lld still active", ticket);
    s->all.erase(it);
    return HHS_OK;
}

int hhs_server_stats(hhs_server *s, hhs_server_stats_t *out) {
    HHS_REQUIRE(s && out, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->now = s->now();
    out->used_bytes = s->used;
    out->peak_bytes = s->peak;
    out->pending = static_cast<int>(s->pending.size());
    out->running_batches = static_cast<int>(s->inflight.size());
    out->running_big = s->big_running;
    return HHS_OK;
}

}  // extern "C"
