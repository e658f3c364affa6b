// pp_serve.hip -- the tile-serving loop behind the C ABI (host code only).
//
// Reference: run_multitile_split's per-device task loop (src/MultiResolution.jl:
// 1183-1203: one Threads.@spawn task per device pulling tiles round-robin,
// each running process_tile!, :719-1180).  Here one GPU serves a stream of
// independent tiles with continuous batching:
//  * slots: each slot is a worker thread with its own pp_ctx (own HIP streams,
//    pooled buffers, resident tables), so tiles on different slots overlap on
//    the device while each tile's work stays in order on its slot's stream;
//  * admission: FIFO; the oldest waiting tile starts when a slot is idle, the
//    declared footprints of the tiles in flight plus its own fit the budget,
//    and the device's live free memory covers its footprint;
//  * the caller's tile function does the tile's work on the slot's context;
//    a failing tile (an error code, or a fault injected for tests) fails
//    alone and the server keeps serving;
//  * completions are reported per ticket (poll) or in completion order
//    (next).
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "pp_common.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t ticket;
    void *request;
    size_t need;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 1;  // 1 = not finished
    int slot = -1;
    std::string msg;  // pp_last_error() of a failed tile, from its worker thread
};

double now_s(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

}  // namespace

struct pp_server {
    pp_serve_config cfg;
    Clock::time_point t0;
    std::vector<pp_ctx *> ctx;
    std::vector<char> busy;
    std::vector<std::thread> workers;
    std::vector<int64_t> assigned;  // per slot: ticket to run, -1 idle
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Job> jobs;
    std::unordered_set<int64_t> inject;
    std::deque<int64_t> completed;  // completion order, for next()
    int64_t next_ticket = 0;
    size_t reserved = 0, budget = 0;
    int in_flight = 0, max_in_flight = 0;
    int64_t n_done = 0, n_failed = 0;
    bool closing = false;

    // the device's free memory, read at admission, covers the tile's
    // footprint (with nothing in flight a tile is always started: if the
    // memory is truly not there, the tile fails on its own allocation)
    bool live_memory_fits(size_t need)
    {
        if (cfg.device < 0 || in_flight == 0) return true;
        size_t fr = 0, tot = 0;
        if (hipSetDevice(cfg.device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
        return fr >= need;
    }

    // under mu: start as many waiting tiles as slots and memory allow
    void schedule()
    {
        while (!pending.empty()) {
            Job &j = jobs[pending.front()];
            int s = -1;
            for (int q = 0; q < (int)busy.size(); q++)
                if (!busy[q]) { s = q; break; }
            if (s < 0 || reserved + j.need > budget || !live_memory_fits(j.need)) return;
            pending.pop_front();
            busy[s] = 1;
            reserved += j.need;
            in_flight++;
            if (in_flight > max_in_flight) max_in_flight = in_flight;
            j.slot = s;
            j.t_start = now_s(t0);
            assigned[s] = j.ticket;
            cv_sched.notify_all();
        }
    }

    void worker(int s)
    {
        if (cfg.device >= 0) (void)hipSetDevice(cfg.device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_sched.wait(lk, [&] { return assigned[s] >= 0 || (closing && pending.empty()); });
            if (assigned[s] < 0) return;
            const int64_t t = assigned[s];
            Job &j = jobs[t];
            const bool injected = inject.erase(t) > 0;
            void *req = j.request;
            lk.unlock();
            int rc;
            if (injected) {
                pp::set_error("fault injected into tile %lld", (long long)t);
                rc = PP_ERR_INJECTED;
            } else {
                rc = cfg.fn(cfg.user, ctx[s], s, t, req);
            }
            std::string msg = rc != PP_OK ? std::string(pp_last_error()) : std::string();
            lk.lock();
            Job &jd = jobs[t];
            jd.status = rc;
            jd.msg = std::move(msg);
            jd.t_done = now_s(t0);
            assigned[s] = -1;
            busy[s] = 0;
            reserved -= jd.need;
            in_flight--;
            n_done++;
            if (rc != PP_OK) n_failed++;
            completed.push_back(t);
            schedule();
            cv_done.notify_all();
            cv_sched.notify_all();
        }
    }
};

using namespace pp;

extern "C" {

int pp_serve_create(const pp_serve_config *cfg, pp_server **out)
{
    PP_ARG(cfg && out, "NULL argument");
    PP_ARG(cfg->fn, "no tile function");
    PP_ARG(cfg->slots >= 1 && cfg->slots <= 64, "slots %d outside [1, 64]", cfg->slots);
    PP_ARG(cfg->queue_capacity >= 1, "queue capacity must be >= 1");
    PP_ARG(cfg->budget_bytes > 0 || (cfg->hbm_fraction > 0 && cfg->hbm_fraction <= 1 && cfg->device >= 0),
           "give an admission budget, or a device and the fraction of its free memory to use");
    *out = nullptr;
    pp_server *s = new pp_server();
    s->cfg = *cfg;
    s->t0 = Clock::now();
    s->ctx.assign(cfg->slots, nullptr);
    if (cfg->device >= 0) {
        for (int q = 0; q < cfg->slots; q++) {
            int rc = pp_ctx_create(cfg->device, &s->ctx[q]);
            if (rc != PP_OK) {
                for (pp_ctx *c : s->ctx) pp_ctx_destroy(c);
                delete s;
                return rc;
            }
        }
    }
    if (cfg->budget_bytes > 0) {
        s->budget = cfg->budget_bytes;
    } else {
        size_t fr = 0, tot = 0;
        hipError_t e = hipSetDevice(cfg->device);
        if (e == hipSuccess) e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) {
            set_error("hipMemGetInfo failed: %s", hipGetErrorString(e));
            for (pp_ctx *c : s->ctx) pp_ctx_destroy(c);
            delete s;
            return PP_ERR_HIP;
        }
        s->budget = (size_t)(cfg->hbm_fraction * (double)fr);
    }
    s->busy.assign(cfg->slots, 0);
    s->assigned.assign(cfg->slots, -1);
    for (int q = 0; q < cfg->slots; q++) s->workers.emplace_back(&pp_server::worker, s, q);
    *out = s;
    return PP_OK;
}

int pp_serve_submit(pp_server *s, void *request, size_t footprint_bytes, int64_t *ticket)
{
    PP_ARG(s && ticket, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->closing) {
        set_error("server is closing");
        return PP_ERR_STATE;
    }
    if (footprint_bytes > s->budget) {
        set_error("tile needs %zu B of device memory, the admission budget is %zu B", footprint_bytes, s->budget);
        return PP_ERR_CAPACITY;
    }
    if ((int)s->pending.size() >= s->cfg.queue_capacity) {
        set_error("serving queue full (%d waiting)", (int)s->pending.size());
        return PP_ERR_BUSY;
    }
    const int64_t t = s->next_ticket++;
    Job j;
    j.ticket = t;
    j.request = request;
    j.need = footprint_bytes;
    j.t_submit = now_s(s->t0);
    s->jobs.emplace(t, j);
    s->pending.push_back(t);
    *ticket = t;
    s->schedule();
    return PP_OK;
}

int pp_serve_inject_fault(pp_server *s, int64_t ticket)
{
    PP_ARG(s, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    PP_ARG(it != s->jobs.end(), "unknown ticket %lld", (long long)ticket);
    if (it->second.slot >= 0) {
        set_error("tile %lld has already started", (long long)ticket);
        return PP_ERR_STATE;
    }
    s->inject.insert(ticket);
    return PP_OK;
}

static void fill_times(const Job &j, double *times)
{
    if (!times) return;
    times[0] = j.t_submit;
    times[1] = j.t_start;
    times[2] = j.t_done;
    times[3] = (double)j.slot;
}

int pp_serve_poll(pp_server *s, int64_t ticket, int timeout_ms, int *status, double *times)
{
    PP_ARG(s && status, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto done = [&] {
        auto it = s->jobs.find(ticket);
        return it == s->jobs.end() || it->second.status != 1;
    };
    if (s->jobs.find(ticket) == s->jobs.end()) {
        set_error("unknown ticket %lld", (long long)ticket);
        return PP_ERR_ARG;
    }
    if (!done()) {
        if (timeout_ms < 0) s->cv_done.wait(lk, done);
        else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
            set_error("tile %lld not finished after %d ms", (long long)ticket, timeout_ms);
            return PP_ERR_TIMEOUT;
        }
    }
    const Job &j = s->jobs[ticket];
    *status = j.status;
    fill_times(j, times);
    return PP_OK;
}

int pp_serve_next(pp_server *s, int timeout_ms, int64_t *ticket, void **request, int *status, double *times)
{
    PP_ARG(s && ticket && status, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->completed.empty(); };
    if (!ready()) {
        if (timeout_ms < 0) s->cv_done.wait(lk, ready);
        else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
            set_error("no tile finished within %d ms", timeout_ms);
            return PP_ERR_TIMEOUT;
        }
    }
    const int64_t t = s->completed.front();
    s->completed.pop_front();
    const Job &j = s->jobs[t];
    *ticket = t;
    if (request) *request = j.request;
    *status = j.status;
    fill_times(j, times);
    return PP_OK;
}

int pp_serve_message(pp_server *s, int64_t ticket, char *buf, size_t len)
{
    PP_ARG(s && buf && len > 0, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    PP_ARG(it != s->jobs.end(), "unknown ticket %lld", (long long)ticket);
    snprintf(buf, len, "%s", it->second.msg.c_str());
    return PP_OK;
}

int pp_serve_stats(pp_server *s, int64_t *stats)
{
    PP_ARG(s && stats, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = (int64_t)s->pending.size();
    stats[1] = s->in_flight;
    stats[2] = s->max_in_flight;
    stats[3] = s->n_done;
    stats[4] = s->n_failed;
    stats[5] = (int64_t)s->reserved;
    stats[6] = (int64_t)s->budget;
    return PP_OK;
}

int pp_serve_context(pp_server *s, int slot, pp_ctx **ctx)
{
    PP_ARG(s && ctx, "NULL argument");
    PP_ARG(slot >= 0 && slot < (int)s->ctx.size(), "slot %d out of range", slot);
    *ctx = s->ctx[slot];
    return PP_OK;
}

int pp_serve_destroy(pp_server *s)
{
    if (!s) return PP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->closing = true;
        s->cv_sched.notify_all();
    }
    for (auto &w : s->workers) w.join();
    for (pp_ctx *c : s->ctx) pp_ctx_destroy(c);
    delete s;
    return PP_OK;
}

}  // extern "C"
