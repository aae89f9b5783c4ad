// -*- C++ -*-
/// @file serve.cpp
/// @brief Continuous batching of kernel launches over the batch launch
/// (loom_server_* in loom_amd.h).  Host code only, written against the
/// public C ABI of the library: device memory, copies, fills and
/// loom_batch_run.
///
/// The nearest relative in the reference is the Shop's exchange loop: a
/// stream admits new inputs as old threads halt (src/vm/mp.hpp:61-131,
/// 316-385; src/vm/simd.hpp:2797-2903).  Here the unit of admission is a
/// whole request (a launch of n threads): it becomes resident when its
/// device columns fit the HBM budget, and every step advances ALL resident
/// requests by a slice of their threads in ONE persistent batch launch.
#include <loom_amd.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
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
    std::uint64_t      req    = 0;
    std::uint64_t      cursor = 0;
    std::uint64_t      off    = 0, bytes = 0;
    std::uint64_t      ticket = 0; ///< last copy of this request (0: none)
    std::vector<char*> in, out;
};
} // namespace

struct loom_server
{
    loom_server_config cfg{};
    char*              base = nullptr; // the arena's device block
    Arena              arena;
    std::vector<int>   members;        // global kernel indices of the set
};

extern "C" {

int
loom_server_create(const loom_server_config* config, loom_server** server)
{
    if(!config || !server) return LOOM_E_ARG;
    if(config->max_resident < 1 || config->max_resident > LOOM_MAX_JOBS) return LOOM_E_ARG;
    if(config->hbm_budget < kAlign || config->step_threads < 64) return LOOM_E_ARG;
    auto s = new loom_server;
    s->cfg = *config;
    int buf[LOOM_MAX_JOBS];
    const int nm = loom_batch_set_members(config->batch_set, buf, LOOM_MAX_JOBS);
    if(nm <= 0)
    {
        delete s;
        return nm < 0 ? -nm : LOOM_E_KERNEL;
    }
    s->members.assign(buf, buf + std::min(nm, static_cast<int>(LOOM_MAX_JOBS)));
    void* p = nullptr;
    if(int r = loom_malloc(&p, config->hbm_budget))
    {
        delete s;
        return r;
    }
    s->base = static_cast<char*>(p);
    s->arena.reset(config->hbm_budget);
    *server = s;
    return 0;
}

int
loom_server_destroy(loom_server* server)
{
    if(!server) return 0;
    int r = 0;
    if(server->base) r = loom_free(server->base);
    delete server;
    return r;
}

int
loom_server_run(loom_server* s, loom_request* reqs, uint64_t count, int realtime,
                loom_server_stats* stats)
{
    using clock = std::chrono::steady_clock;
    if(!s || (!reqs && count)) return LOOM_E_ARG;
    for(std::uint64_t i = 1; i < count; ++i)
        if(reqs[i].arrival_ms < reqs[i - 1].arrival_ms) return LOOM_E_ARG;

    loom_server_stats st{};
    const auto        t0  = clock::now();
    auto              now = [&] { return std::chrono::duration<double, std::milli>(clock::now() - t0).count(); };

    // kernel infos of the members
    std::map<int, loom_kernel_info> info;
    for(int k: s->members)
    {
        loom_kernel_info ki{};
        if(int r = loom_kernel_info_get(k, &ki)) return r;
        info.emplace(k, ki);
    }
    for(std::uint64_t i = 0; i < count; ++i)
    {
        reqs[i].status   = 0;
        reqs[i].steps    = 0;
        reqs[i].admit_ms = reqs[i].done_ms = 0;
        if(!info.count(reqs[i].kernel)) reqs[i].status = LOOM_E_KERNEL;
        else if(reqs[i].n > 0xFFFFFFFFull) reqs[i].status = LOOM_E_CAPACITY;
    }

    // A request moves through: pending (arrived, waiting for a slot and
    // memory) -> loading (inputs copying in on the copy stream) -> running
    // (in the steps) -> draining (outputs copying out) -> done.  Copies of
    // one request overlap the steps of the others.
    std::deque<std::uint64_t> pending;
    std::vector<Resident>     loading, running, draining;
    std::uint64_t             next = 0, finished = 0;
    std::uint64_t             last_ticket = 0; // latest copy issued by this run

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

    // start copying the outputs of a request whose threads all ran
    auto drain = [&](Resident& r) -> int
    {
        loom_request&           q  = reqs[r.req];
        const loom_kernel_info& ki = info.at(q.kernel);
        r.ticket                   = 0;
        for(int k = 0; k < ki.n_outputs && q.n; ++k)
        {
            if(int e = loom_copy_async(q.outputs[k], r.out[k], q.n * type_bytes(ki.output_type[k]),
                                       LOOM_COPY_D2H, &r.ticket))
                return e;
            last_ticket = r.ticket;
        }
        draining.push_back(std::move(r));
        return 0;
    };

    auto ready = [&](const Resident& r, int& done) -> int
    {
        done = 1;
        return r.ticket ? loom_copy_status(r.ticket, &done) : 0;
    };

    // on an error, let the copies already issued land before the arena is
    // handed to another run, then give the whole arena back
    auto fail = [&](int e) -> int
    {
        if(last_ticket) loom_copy_wait(last_ticket);
        loom_synchronize();
        s->arena.reset(s->cfg.hbm_budget);
        return e;
    };

    std::vector<std::uint64_t> cols;
    std::vector<loom_job>      jobs;
    while(finished < count)
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
            loom_request& q = reqs[draining[i].req];
            s->arena.release(draining[i].off, draining[i].bytes);
            q.done_ms = now();
            st.threads += q.n;
            ++finished;
            draining.erase(draining.begin() + static_cast<std::ptrdiff_t>(i));
        }
        // ---- arrivals ----
        const double t = now();
        while(next < count && (!realtime || reqs[next].arrival_ms <= t))
        {
            if(reqs[next].status)
                ++finished; // rejected before admission
            else
                pending.push_back(next);
            ++next;
        }
        // ---- admission: FIFO, by slot and HBM budget; inputs start copying ----
        while(!pending.empty() && loading.size() + running.size() < s->cfg.max_resident)
        {
            loom_request&       q     = reqs[pending.front()];
            const std::uint64_t bytes = request_bytes(q, cols);
            if(bytes > s->cfg.hbm_budget)
            {
                q.status = LOOM_E_CAPACITY;
                pending.pop_front();
                ++finished;
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
                                           loom_fill(r.in[k], &q.scalars[k], tb, q.n) :
                                           loom_copy_async(r.in[k], q.inputs[k], q.n * tb,
                                                           LOOM_COPY_H2D, &r.ticket);
                if(e) return fail(e);
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
                if(int e = loom_copy_wait(loading.front().ticket)) return fail(e);
            }
            else if(!draining.empty())
            {
                if(int e = loom_copy_wait(draining.front().ticket)) return fail(e);
            }
            else if(next < count && realtime)
            {
                const double wait = reqs[next].arrival_ms - now();
                if(wait > 0.05)
                    std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait - 0.02));
            }
            continue;
        }
        // ---- one step: a slice of every running request in one batch launch ----
        const std::uint64_t quota =
            std::max<std::uint64_t>(64, (s->cfg.step_threads / running.size() + 63) / 64 * 64);
        jobs.assign(running.size(), loom_job{});
        for(std::size_t i = 0; i < running.size(); ++i)
        {
            const loom_request& q = reqs[running[i].req];
            loom_job&           j = jobs[i];
            j.kernel              = q.kernel;
            j.begin               = running[i].cursor;
            j.end                 = std::min<std::uint64_t>(q.n, running[i].cursor + quota);
            for(std::size_t k = 0; k < running[i].in.size(); ++k) j.inputs[k] = running[i].in[k];
            for(std::size_t k = 0; k < running[i].out.size(); ++k) j.outputs[k] = running[i].out[k];
        }
        loom_run_stats rs{};
        if(int e = loom_batch_run(s->cfg.batch_set, jobs.data(), static_cast<int>(jobs.size()),
                                  &s->cfg.params, &rs))
        {
            for(auto& r: running) reqs[r.req].status = e;
            return fail(e);
        }
        st.gpu_ms += rs.ms_main;
        ++st.steps;
        std::vector<Resident> keep;
        keep.reserve(running.size());
        for(std::size_t i = 0; i < running.size(); ++i)
        {
            running[i].cursor = jobs[i].end;
            ++reqs[running[i].req].steps;
            if(running[i].cursor >= reqs[running[i].req].n)
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

} // extern "C"
