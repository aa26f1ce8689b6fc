// Continuous-batching server over the CS enumeration engine.
//
// A worker thread repeats: admit queued requests (slots and HBM budget
// permitting), then one engine step, which builds the next size level of every
// running request in shared launches -- requests at different levels share the
// same k_expand / k_resolve / k_store launches.  Finished requests are handed
// to their waiters; a request that could not get arena memory is restarted
// when memory has freed up, so every result equals a plain solve.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "engine.h"

namespace vfb {
namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    vfb_instance inst;
    vfb_result res{};
    Inst st;
    bool done = false;
    Clock::time_point t_submit;
};

void reset_req(Req& r, const vfb_options& opt, uint64_t budget) {
    std::free(r.res.formula);
    const uint32_t restarts = r.st.deferred ? r.res.restarts + 1 : 0;
    std::memset(&r.res, 0, sizeof r.res);
    r.res.status = VFB_NOT_FOUND;
    r.res.restarts = restarts;
    r.st = Inst{};
    r.st.in = &r.inst;
    r.st.res = &r.res;
    r.st.max_size = opt.max_size;
    r.st.max_rows = opt.max_rows;
    r.st.budget = budget;
    r.st.prio = r.ticket;
}

}  // namespace
}  // namespace vfb

using namespace vfb;

struct vfb_server {
    vfb_options opt;
    uint64_t budget = 0;        // admission bound on the memory of running requests
    uint64_t per_req = 0;       // budget of one request (as vfb_solve)
    int max_active = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    // requests that could not get arena memory beside the others: each runs
    // alone, after the running ones finish, as its plain solve would
    std::deque<Req*> solo;
    Req* solo_run = nullptr;
    std::map<uint64_t, std::unique_ptr<Req>> reqs;
    std::vector<Req*> active;
    uint64_t next_ticket = 1;
    uint64_t peak_active = 0;
    uint64_t bytes_active = 0;
    bool stop = false;
    int waiters = 0;            // callers blocked in vfb_server_wait
    std::thread th;
    void loop();
};

static int g_servers = 0;   // guarded by engine_mutex(): an engine in use by a server is never re-created

void vfb_server::loop() {
    for (;;) {
        std::vector<Req*> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty() || !solo.empty(); });
            if (stop) return;
            uint64_t used = 0;
            for (Req* r : active) used += r->st.used;
            if (!solo_run && !solo.empty() && active.empty()) {
                solo_run = solo.front();   // the arena is empty now: it runs alone
                solo.pop_front();
                admit.push_back(solo_run);
            }
            while (!solo_run && solo.empty() && !queue.empty() && (int)(active.size() + admit.size()) < max_active &&
                   used < budget) {
                admit.push_back(queue.front());
                queue.pop_front();
                used += 1 << 20;   // a new request starts with about this much (atoms, tables)
            }
            bytes_active = used;
        }
        std::vector<Req*> touched = active;
        touched.insert(touched.end(), admit.begin(), admit.end());
        {
            std::lock_guard<std::mutex> g(engine_mutex());
            try {
                Engine& eng = engine_acquire(0);
                eng.begin(opt, true);
                eng.set_defer_alone(solo_run == nullptr);
                std::vector<Inst*> fresh, act;
                for (Req* r : admit) {
                    reset_req(*r, opt, per_req);
                    fresh.push_back(&r->st);
                }
                for (Req* r : active) act.push_back(&r->st);
                eng.start(fresh, act);
                if (!act.empty()) eng.step(act);
            } catch (std::exception& e) {
                std::fprintf(stderr, "vfb server: device error: %s\n", e.what());
                // other servers' requests may live in the engine too: release
                // only this server's running and admitted requests
                std::vector<Inst*> mine;
                for (Req* r : touched) mine.push_back(&r->st);
                try {
                    engine_acquire(0).forget(mine);
                } catch (std::exception&) {
                }
                for (Req* r : touched) {
                    r->st.done = true;
                    r->res.status = VFB_DEVICE_ERROR;
                    if (!r->res.formula) r->res.formula = strdup("see_the_error");
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            active.clear();
            const auto now = Clock::now();
            for (Req* r : touched) {
                if (r == solo_run && r->st.done) solo_run = nullptr;
                if (!r->st.done) {
                    active.push_back(r);
                } else if (r->st.deferred) {
                    if (r->st.need_alone) solo.push_back(r);   // restarted alone
                    else queue.push_front(r);                  // restarted once memory is free
                } else {
                    r->res.ms_total = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
                    if (!r->res.formula) r->res.formula = strdup("Not found !");
                    r->done = true;
                }
            }
            peak_active = std::max<uint64_t>(peak_active, active.size());
        }
        cv_done.notify_all();
    }
}

extern "C" vfb_status vfb_server_create(const vfb_options* opt, uint64_t hbm_budget, int max_active,
                                        vfb_server** out) {
    if (!out) return VFB_INVALID_INPUT;
    *out = nullptr;
    auto s = std::make_unique<vfb_server>();
    vfb_options_default(&s->opt);
    if (opt) s->opt = *opt;
    if (s->opt.relaxed_type < 1 || s->opt.relaxed_type > 3) s->opt.relaxed_type = 3;
    if (s->opt.max_size < 1 || s->opt.max_size > VFB_MAX_SIZE) return VFB_INVALID_INPUT;
    // CS levels of unit cost only
    if (s->opt.search != VFB_SEARCH_CS || s->opt.bs_max_size) return VFB_INVALID_INPUT;
    for (int k = 0; k < 8; ++k)
        if (s->opt.cost[k] > 1) return VFB_INVALID_INPUT;
    s->max_active = max_active > 0 ? max_active : 1024;
    {
        std::lock_guard<std::mutex> g(engine_mutex());
        try {
            Engine& eng = engine_acquire(s->opt.hbm_budget);
            const uint64_t arena = eng.arena_size();
            s->per_req = s->opt.hbm_budget ? std::min<uint64_t>(s->opt.hbm_budget, arena) : arena;
            s->budget = hbm_budget ? std::min<uint64_t>(hbm_budget, arena) : arena;
            ++g_servers;
        } catch (std::exception& e) {
            std::fprintf(stderr, "vfb server: device error: %s\n", e.what());
            return VFB_DEVICE_ERROR;
        }
    }
    vfb_server* raw = s.get();
    s->th = std::thread([raw] { raw->loop(); });
    *out = s.release();
    return VFB_OK;
}

extern "C" vfb_status vfb_server_submit(vfb_server* s, const vfb_instance* inst, uint64_t* ticket) {
    if (!s || !inst) return VFB_INVALID_INPUT;
    std::string why;
    if (validate(*inst, &why) != VFB_OK) return VFB_INVALID_INPUT;
    auto r = std::make_unique<Req>();
    r->inst = *inst;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        if (ticket) *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->reqs[r->ticket] = std::move(r);
    }
    s->cv_work.notify_one();
    return VFB_OK;
}

// Hands a finished request's result to its caller and releases the ticket.
// A ticket is consumed once: a second caller waiting on the same ticket, or
// one still waiting when the server stops, gets VFB_INVALID_INPUT.  The
// request is looked up again after every wake-up, since another caller may
// have consumed (and freed) it meanwhile.
static vfb_status take(vfb_server* s, uint64_t ticket, vfb_result* out, bool block, int* done) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (done) *done = 0;
    if (s->reqs.find(ticket) == s->reqs.end()) return VFB_INVALID_INPUT;
    if (block) {
        ++s->waiters;
        s->cv_done.wait(lk, [&] {
            auto w = s->reqs.find(ticket);
            return s->stop || w == s->reqs.end() || w->second->done;
        });
        --s->waiters;
        if (s->stop) s->cv_done.notify_all();   // vfb_server_destroy waits for the waiters to leave
    }
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return VFB_INVALID_INPUT;
    Req* r = it->second.get();
    if (!r->done) return block ? VFB_INVALID_INPUT : VFB_OK;   // blocked, and the server stopped
    if (done) *done = 1;
    if (out) {
        *out = r->res;
        r->res.formula = nullptr;
    }
    const vfb_status st = r->res.status;
    std::free(r->res.formula);
    s->reqs.erase(it);
    return st;
}

extern "C" vfb_status vfb_server_poll(vfb_server* s, uint64_t ticket, vfb_result* out, int* done) {
    if (!s) return VFB_INVALID_INPUT;
    return take(s, ticket, out, false, done);
}

extern "C" vfb_status vfb_server_wait(vfb_server* s, uint64_t ticket, vfb_result* out) {
    if (!s) return VFB_INVALID_INPUT;
    return take(s, ticket, out, true, nullptr);
}

extern "C" void vfb_server_stats(vfb_server* s, uint64_t* queued, uint64_t* active, uint64_t* bytes,
                                 uint64_t* peak_active) {
    if (!s) return;
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = s->queue.size() + s->solo.size();
    if (active) *active = s->active.size();
    if (bytes) *bytes = s->bytes_active;
    if (peak_active) *peak_active = s->peak_active;
}

extern "C" void vfb_server_destroy(vfb_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    {
        // callers blocked in vfb_server_wait return VFB_INVALID_INPUT before
        // the server (its mutex and condition variables) is freed
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] { return s->waiters == 0; });
    }
    {
        std::lock_guard<std::mutex> g(engine_mutex());
        // release the device memory of requests still running
        try {
            Engine& eng = engine_acquire(0);
            eng.begin(s->opt, true);
            for (Req* r : s->active) eng.abandon(r->st);
            eng.drain();
        } catch (std::exception&) {
        }
        --g_servers;
    }
    for (auto& kv : s->reqs) std::free(kv.second->res.formula);
    delete s;
}

namespace vfb {
bool engine_pinned() { return g_servers > 0; }
}  // namespace vfb
