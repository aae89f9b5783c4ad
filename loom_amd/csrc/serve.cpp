// -*- C++ -*-
/// @file serve.cpp
/// @brief Continuous batching of kernel launches over the batch launch
/// (loom_server_* in loom_amd.h).  Host code only, written against the
/// public C ABI of the library: device memory, and copies, fills and batch
/// launches on an execution context of the server's own.
///
/// The nearest relative in the reference is the Shop's exchange loop: a
/// stream admits new inputs as old threads halt (src/vm/mp.hpp:61-131,
/// 316-385; src/vm/simd.hpp:2797-2903).  Here the unit of admission is a
/// whole request (a launch of n threads): it becomes resident when its
/// device columns fit the HBM budget, and every step advances ALL resident
/// requests by a slice of their threads in ONE persistent batch launch.
///
/// Requests come from a source: a closed trace with arrival times
/// (loom_server_run), or submissions from any host thread while the loop
/// runs on a thread of its own (loom_server_start / submit / poll / stop:
/// the Shop's main thread exchanging while its workers compute,
/// src/vm/mp.hpp:349-369).
#include <loom_amd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <mutex>
#include <thread>
#include <vector>

namespace
{
constexpr std::uint64_t kAlign = 256;

std::size_t
type_bytes(int t)
{
    switch(t)
    {
    case LOOM_U8:
    case LOOM_I8: return 1;
    case LOOM_U16:
    case LOOM_I16: return 2;
    case LOOM_U32:
    case LOOM_I32:
    case LOOM_F32: return 4;
    default: return 8;
    }
}

/// First-fit allocator over one device block (offsets, host bookkeeping).
class Arena
{
    std::map<std::uint64_t, std::uint64_t> m_free; // offset -> size
    std::uint64_t                          m_used = 0;

  public:
    explicit Arena(std::uint64_t size = 0) { reset(size); }

    void
    reset(std::uint64_t size)
    {
        m_free.clear();
        if(size) m_free.emplace(0, size);
        m_used = 0;
    }

    bool
    alloc(std::uint64_t bytes, std::uint64_t& off)
    {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        for(auto it = m_free.begin(); it != m_free.end(); ++it)
            if(it->second >= bytes)
            {
                off                    = it->first;
                const std::uint64_t sz = it->second;
                m_free.erase(it);
                if(sz > bytes) m_free.emplace(off + bytes, sz - bytes);
                m_used += bytes;
                return true;
            }
        return false;
    }

    void
    release(std::uint64_t off, std::uint64_t bytes)
    {
        bytes = (bytes + kAlign - 1) / kAlign * kAlign;
        m_used -= bytes;
        auto it = m_free.emplace(off, bytes).first;
        // merge with the next and the previous free range
        auto nx = std::next(it);
        if(nx != m_free.end() && it->first + it->second == nx->first)
        {
            it->second += nx->second;
            m_free.erase(nx);
        }
        if(it != m_free.begin())
        {
            auto pv = std::prev(it);
            if(pv->first + pv->second == it->first)
            {
                pv->second += it->second;
                m_free.erase(it);
            }
        }
    }

    std::uint64_t
    used() const
    {
        return m_used;
    }
};

struct Resident
{
    loom_request*      req    = nullptr;
    std::uint64_t      cursor = 0;
    std::uint64_t      off    = 0, bytes = 0;
    std::uint64_t      ticket = 0; ///< last copy of this request (0: none)
    std::vector<char*> in, out;
};
} // namespace


namespace
{
using clock_type = std::chrono::steady_clock;

/// @brief Where requests come from.  `take` moves the requests that have
/// arrived by `now` (ms since the loop started) into `out` and returns
/// false once no request will ever come again; `idle` blocks until one may
/// arrive (at most `max_ms`).
struct Source
{
    virtual ~Source()                                                  = default;
    virtual bool take(double now, std::vector<loom_request*>& out)     = 0;
    virtual void idle(double now, double max_ms)                       = 0;
};

/// A closed trace sorted by arrival time, replayed against the clock
/// (realtime) or all arrived at the start.
struct TraceSource final : Source
{
    loom_request* reqs;
    std::uint64_t count, next = 0;
    bool          realtime;

    TraceSource(loom_request* r, std::uint64_t c, bool rt) : reqs(r), count(c), realtime(rt) {}

    bool
    take(double now, std::vector<loom_request*>& out) override
    {
        while(next < count && (!realtime || reqs[next].arrival_ms <= now)) out.push_back(&reqs[next++]);
        return next < count;
    }

    void
    idle(double now, double max_ms) override
    {
        if(next >= count || !realtime) return;
        const double wait = std::min(reqs[next].arrival_ms - now, max_ms);
        if(wait > 0.05)
            std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait - 0.02));
    }
};

/// Submissions from other host threads (a locked queue the loop drains).
struct OnlineSource final : Source
{
    std::mutex                mu;
    std::condition_variable   cv;
    std::deque<loom_request*> queue;
    bool                      open     = false; ///< submissions accepted
    bool                      stopping = false;

    bool
    take(double, std::vector<loom_request*>& out) override
    {
        std::lock_guard<std::mutex> lk(mu);
        while(!queue.empty())
        {
            out.push_back(queue.front());
            queue.pop_front();
        }
        return !stopping;
    }

    void
    idle(double, double max_ms) override
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait_for(lk, std::chrono::duration<double, std::milli>(max_ms),
                    [&] { return !queue.empty() || stopping; });
    }
};
} // namespace

struct loom_server
{
    loom_server_config cfg{};
    loom_context*      ctx  = nullptr; // the server's own kernel and copy streams
    char*              base = nullptr; // the arena's device block
    Arena              arena;
    std::vector<int>   members;        // global kernel indices of the set
    // online serving
    std::mutex             mu;          // start / stop
    OnlineSource           online;
    std::thread            loop;
    bool                   running = false;
    clock_type::time_point t0{};
    int                    loop_status = 0;
    loom_server_stats      loop_stats{};
};

namespace
{
void
mark_complete(loom_request& q)
{
    __atomic_store_n(&q.complete, 1, __ATOMIC_RELEASE);
}

/// @brief The serving loop over requests from @p src until it is exhausted
/// and every request it gave has completed.
int
serve(loom_server* s, Source& src, clock_type::time_point t0, loom_server_stats* stats)
{
    loom_context*     ctx = s->ctx;
    loom_server_stats st{};
    auto now = [&] { return std::chrono::duration<double, std::milli>(clock_type::now() - t0).count(); };

    // kernel infos: every registered kernel can be served; members of the
    // batch set run together in one batch launch per step, any other kernel
    // runs its slice of the step as a launch of its own
    std::map<int, loom_kernel_info> info;
    for(int k = 0, nk = loom_kernel_count(); k < nk; ++k)
    {
        loom_kernel_info ki{};
        if(int r = loom_kernel_info_get(k, &ki)) return r;
        info.emplace(k, ki);
    }
    const std::set<int> in_set(s->members.begin(), s->members.end());

    // A request moves through: pending (arrived, waiting for a slot and
    // memory) -> loading (inputs copying in on the copy stream) -> running
    // (in the steps) -> draining (outputs copying out) -> done.  Copies of
    // one request overlap the steps of the others.
    std::deque<loom_request*>  pending;
    std::vector<Resident>      loading, running, draining;
    std::vector<loom_request*> arrived;
    std::uint64_t              last_ticket = 0; // latest copy issued by this loop
    bool                       more        = true;

    auto request_bytes = [&](const loom_request& q, std::vector<std::uint64_t>& cols)
    {
        const loom_kernel_info& ki = info.at(q.kernel);
        std::uint64_t           b  = 0;
        cols.clear();
        for(int k = 0; k < ki.n_inputs + ki.n_outputs; ++k)
        {
            const int           t = k < ki.n_inputs ? ki.input_type[k] : ki.output_type[k - ki.n_inputs];
            const std::uint64_t c = (std::max<std::uint64_t>(q.n, 1) * type_bytes(t) + kAlign - 1) /
                                    kAlign * kAlign;
            cols.push_back(c);
            b += c;
        }
        return b;
    };

    auto finish = [&](loom_request& q)
    {
        q.done_ms = now();
        if(q.status == 0) st.threads += q.n;
        mark_complete(q);
    };

    // start copying the outputs of a request whose threads all ran
    auto drain = [&](Resident& r) -> int
    {
        loom_request&           q  = *r.req;
        const loom_kernel_info& ki = info.at(q.kernel);
        r.ticket                   = 0;
        for(int k = 0; k < ki.n_outputs && q.n; ++k)
        {
            if(int e = loom_ctx_copy_async(ctx, q.outputs[k], r.out[k],
                                           q.n * type_bytes(ki.output_type[k]), LOOM_COPY_D2H,
                                           &r.ticket))
                return e;
            last_ticket = r.ticket;
        }
        draining.push_back(std::move(r));
        return 0;
    };

    auto ready = [&](const Resident& r, int& done) -> int
    {
        done = 1;
        return r.ticket ? loom_ctx_copy_status(ctx, r.ticket, &done) : 0;
    };

    // on an error, let the copies already issued land before the arena is
    // reused, give the whole arena back, and complete every request the
    // loop holds with the error
    auto fail = [&](int e) -> int
    {
        if(last_ticket) loom_ctx_copy_wait(ctx, last_ticket);
        loom_ctx_synchronize(ctx);
        s->arena.reset(s->cfg.hbm_budget);
        for(auto* v: {&loading, &running, &draining})
            for(auto& r: *v)
            {
                if(!r.req->status) r.req->status = e;
                finish(*r.req);
            }
        for(auto* q: pending)
        {
            q->status = e;
            finish(*q);
        }
        // requests still to come are refused with the same error
        for(;;)
        {
            arrived.clear();
            const bool again = src.take(1e300, arrived);
            for(auto* q: arrived)
            {
                q->status = e;
                finish(*q);
            }
            if(!again) break;
            src.idle(now(), 1.0);
        }
        return e;
    };

    auto admit_new = [&](loom_request* q)
    {
        q->status   = 0;
        q->steps    = 0;
        q->admit_ms = q->done_ms = 0;
        if(!info.count(q->kernel))
            q->status = LOOM_E_KERNEL;
        else if(q->n > 0xFFFFFFFFull)
            q->status = LOOM_E_CAPACITY;
        if(q->status)
            finish(*q); // rejected before admission
        else
            pending.push_back(q);
    };

    std::vector<std::uint64_t> cols;
    std::vector<loom_job>      jobs;
    for(;;)
    {
        // ---- retire requests whose outputs arrived on the host ----
        for(std::size_t i = 0; i < draining.size();)
        {
            int done = 0;
            if(int e = ready(draining[i], done)) return fail(e);
            if(!done)
            {
                ++i;
                continue;
            }
            s->arena.release(draining[i].off, draining[i].bytes);
            finish(*draining[i].req);
            draining.erase(draining.begin() + static_cast<std::ptrdiff_t>(i));
        }
        // ---- arrivals ----
        arrived.clear();
        more = src.take(now(), arrived);
        for(auto* q: arrived) admit_new(q);
        if(!more && pending.empty() && loading.empty() && running.empty() && draining.empty())
            break;
        // ---- admission: FIFO, by slot and HBM budget; inputs start copying ----
        while(!pending.empty() && loading.size() + running.size() < s->cfg.max_resident)
        {
            loom_request&       q     = *pending.front();
            const std::uint64_t bytes = request_bytes(q, cols);
            if(bytes > s->cfg.hbm_budget)
            {
                q.status = LOOM_E_CAPACITY;
                pending.pop_front();
                finish(q);
                continue;
            }
            std::uint64_t off = 0;
            if(!s->arena.alloc(bytes, off)) break;
            Resident r;
            r.req   = pending.front();
            r.off   = off;
            r.bytes = bytes;
            const loom_kernel_info& ki = info.at(q.kernel);
            char*                   p  = s->base + off;
            for(int k = 0; k < ki.n_inputs + ki.n_outputs; ++k)
            {
                (k < ki.n_inputs ? r.in : r.out).push_back(p);
                p += cols[k];
            }
            for(int k = 0; k < ki.n_inputs && q.n; ++k)
            {
                const std::size_t tb = type_bytes(ki.input_type[k]);
                const int         e  = (q.broadcast >> k) & 1u ?
                                           loom_ctx_fill(ctx, r.in[k], &q.scalars[k], tb, q.n) :
                                           loom_ctx_copy_async(ctx, r.in[k], q.inputs[k], q.n * tb,
                                                               LOOM_COPY_H2D, &r.ticket);
                if(e)
                {
                    running.push_back(std::move(r)); // completed with the error by fail
                    pending.pop_front();
                    return fail(e);
                }
                if(r.ticket) last_ticket = r.ticket;
            }
            q.admit_ms = now();
            pending.pop_front();
            st.peak_bytes = std::max(st.peak_bytes, s->arena.used());
            if(q.n == 0)
            {
                if(int e = drain(r)) return fail(e);
                continue;
            }
            loading.push_back(std::move(r));
        }
        // ---- requests whose inputs arrived join the steps ----
        for(std::size_t i = 0; i < loading.size();)
        {
            int done = 0;
            if(int e = ready(loading[i], done)) return fail(e);
            if(!done)
            {
                ++i;
                continue;
            }
            running.push_back(std::move(loading[i]));
            loading.erase(loading.begin() + static_cast<std::ptrdiff_t>(i));
        }
        st.peak_resident = std::max<std::uint32_t>(
            st.peak_resident, static_cast<std::uint32_t>(loading.size() + running.size()));
        if(running.empty())
        {
            // nothing to run: wait for a copy, or idle until the next arrival
            if(!loading.empty())
            {
                if(int e = loom_ctx_copy_wait(ctx, loading.front().ticket)) return fail(e);
            }
            else if(!draining.empty())
            {
                if(int e = loom_ctx_copy_wait(ctx, draining.front().ticket)) return fail(e);
            }
            else if(more)
                src.idle(now(), 50.0);
            continue;
        }
        // ---- one step: a slice of every running request in one batch launch ----
        const std::uint64_t quota =
            std::max<std::uint64_t>(64, (s->cfg.step_threads / running.size() + 63) / 64 * 64);
        jobs.clear();
        std::vector<std::uint64_t> ends(running.size());
        for(std::size_t i = 0; i < running.size(); ++i)
        {
            const loom_request& q = *running[i].req;
            ends[i] = std::min<std::uint64_t>(q.n, running[i].cursor + quota);
            if(!in_set.count(q.kernel)) continue;
            loom_job j{};
            j.kernel = q.kernel;
            j.begin  = running[i].cursor;
            j.end    = ends[i];
            for(std::size_t k = 0; k < running[i].in.size(); ++k) j.inputs[k] = running[i].in[k];
            for(std::size_t k = 0; k < running[i].out.size(); ++k) j.outputs[k] = running[i].out[k];
            jobs.push_back(j);
        }
        loom_run_stats rs{};
        if(!jobs.empty())
        {
            if(int e = loom_ctx_batch_run(ctx, s->cfg.batch_set, jobs.data(),
                                          static_cast<int>(jobs.size()), &s->cfg.params, &rs))
            {
                for(auto& r: running) r.req->status = e;
                return fail(e);
            }
            st.gpu_ms += rs.ms_main;
        }
        // kernels outside the batch set: their slice of the step, one launch
        // each, on the same stream (columns offset to the slice's first thread)
        for(std::size_t i = 0; i < running.size(); ++i)
        {
            const loom_request& q = *running[i].req;
            if(in_set.count(q.kernel)) continue;
            const loom_kernel_info& ki = info.at(q.kernel);
            const std::uint64_t     b  = running[i].cursor;
            const void*             ins[LOOM_MAX_IO]{};
            void*                   outs[LOOM_MAX_IO]{};
            for(int k = 0; k < ki.n_inputs; ++k)
                ins[k] = running[i].in[k] + b * type_bytes(ki.input_type[k]);
            for(int k = 0; k < ki.n_outputs; ++k)
                outs[k] = running[i].out[k] + b * type_bytes(ki.output_type[k]);
            loom_run_stats ks{};
            if(int e = loom_ctx_run(ctx, q.kernel, LOOM_MODE_HYBRID, ins, outs, ends[i] - b,
                                    &s->cfg.params, &ks))
            {
                for(auto& r: running) r.req->status = e;
                return fail(e);
            }
            st.gpu_ms += ks.ms_main;
        }
        ++st.steps;
        std::vector<Resident> keep;
        keep.reserve(running.size());
        for(std::size_t i = 0; i < running.size(); ++i)
        {
            running[i].cursor = ends[i];
            ++running[i].req->steps;
            if(running[i].cursor >= running[i].req->n)
            {
                if(int e = drain(running[i])) return fail(e);
            }
            else
                keep.push_back(std::move(running[i]));
        }
        running.swap(keep);
    }
    st.makespan_ms = now();
    if(stats) *stats = st;
    return 0;
}
} // namespace

extern "C" {

int
loom_server_create(const loom_server_config* config, loom_server** server)
{
    if(!config || !server) return LOOM_E_ARG;
    if(config->max_resident < 1 || config->max_resident > LOOM_MAX_JOBS) return LOOM_E_ARG;
    if(config->hbm_budget < kAlign || config->step_threads < 64) return LOOM_E_ARG;
    auto s = std::make_unique<loom_server>();
    s->cfg = *config;
    int buf[LOOM_MAX_JOBS];
    const int nm = loom_batch_set_members(config->batch_set, buf, LOOM_MAX_JOBS);
    if(nm <= 0) return nm < 0 ? -nm : LOOM_E_KERNEL;
    s->members.assign(buf, buf + std::min(nm, static_cast<int>(LOOM_MAX_JOBS)));
    if(int r = loom_context_create(&s->ctx)) return r;
    void* p = nullptr;
    if(int r = loom_malloc(&p, config->hbm_budget))
    {
        loom_context_destroy(s->ctx);
        return r;
    }
    s->base = static_cast<char*>(p);
    s->arena.reset(config->hbm_budget);
    *server = s.release();
    return 0;
}

int
loom_server_destroy(loom_server* server)
{
    if(!server) return 0;
    loom_server_stop(server, nullptr);
    int r = 0;
    if(server->ctx) loom_context_destroy(server->ctx);
    if(server->base) r = loom_free(server->base);
    delete server;
    return r;
}

int
loom_server_run(loom_server* s, loom_request* reqs, uint64_t count, int realtime,
                loom_server_stats* stats)
{
    if(!s || (!reqs && count)) return LOOM_E_ARG;
    for(std::uint64_t i = 1; i < count; ++i)
        if(reqs[i].arrival_ms < reqs[i - 1].arrival_ms) return LOOM_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if(s->running) return LOOM_E_STATE;
        s->running = true;
    }
    for(std::uint64_t i = 0; i < count; ++i) reqs[i].complete = 0;
    TraceSource src(reqs, count, realtime != 0);
    const int   r = serve(s, src, clock_type::now(), stats);
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = false;
    return r;
}

int
loom_server_start(loom_server* s)
{
    if(!s) return LOOM_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if(s->running) return LOOM_E_STATE;
    {
        std::lock_guard<std::mutex> q(s->online.mu);
        s->online.stopping = false;
        s->online.queue.clear();
    }
    s->running     = true;
    s->loop_status = 0;
    s->loop_stats  = loom_server_stats{};
    s->t0          = clock_type::now();
    s->loop = std::thread([s] { s->loop_status = serve(s, s->online, s->t0, &s->loop_stats); });
    std::lock_guard<std::mutex> q(s->online.mu);
    s->online.open = true;
    return 0;
}

int
loom_server_submit(loom_server* s, loom_request* q)
{
    if(!s || !q) return LOOM_E_ARG;
    std::lock_guard<std::mutex> lk(s->online.mu);
    if(!s->online.open || s->online.stopping) return LOOM_E_STATE;
    q->complete   = 0;
    q->status     = 0;
    q->arrival_ms = std::chrono::duration<double, std::milli>(clock_type::now() - s->t0).count();
    s->online.queue.push_back(q);
    s->online.cv.notify_one();
    return 0;
}

int
loom_server_poll(loom_server* s, const loom_request* q, int* done)
{
    if(!s || !q || !done) return LOOM_E_ARG;
    *done = __atomic_load_n(&q->complete, __ATOMIC_ACQUIRE) != 0;
    return 0;
}

int
loom_server_wait(loom_server* s, const loom_request* q)
{
    if(!s || !q) return LOOM_E_ARG;
    while(!__atomic_load_n(&q->complete, __ATOMIC_ACQUIRE))
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    return q->status;
}

int
loom_server_stop(loom_server* s, loom_server_stats* stats)
{
    if(!s) return LOOM_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if(!s->loop.joinable()) return 0;
    {
        std::lock_guard<std::mutex> q(s->online.mu);
        s->online.open     = false;
        s->online.stopping = true;
        s->online.cv.notify_one();
    }
    s->loop.join();
    s->running = false;
    if(stats) *stats = s->loop_stats;
    return s->loop_status;
}

} // extern "C"
