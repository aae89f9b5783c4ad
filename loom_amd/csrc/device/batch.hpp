// -*- C++ -*-
/// @file batch.hpp
/// @brief Many launches of different kernels and uneven thread counts in one
/// persistent gfx950 kernel.
///
/// The reference runs several kernels of uneven sizes concurrently with its
/// multi-stream Shop: CPU worker threads round-robin over a bench of
/// per-kernel streams (mp.hpp:158-190, 316-385).  Here one persistent launch
/// takes a table of jobs (kernel, thread range, input/output columns) whose
/// kernels all belong to one compiled batch set.  A wave claims guided
/// chunks of one job, runs them with in-wave refill from the same job, and
/// moves to another job when that one is exhausted; waves start on
/// different jobs (spread by workgroup index) so small and large jobs run
/// side by side.  The register file of a wave is the union of the member
/// kernels' (the largest sets the occupancy).
#pragma once

#include "sched.hpp"

namespace loom::hipvm
{
inline constexpr std::uint32_t kMaxJobs = 64;

/// @brief One job of a batch (device copy of loom_job).
struct JobDesc
{
    std::uint32_t member = 0; ///< index of the kernel in the batch set
    std::uint32_t pad    = 0;
    std::uint64_t begin = 0, end = 0;
    const void*   in[kMaxIO]{};
    void*         out[kMaxIO]{};
};

struct BatchCtl
{
    Slot64 next[kMaxJobs]; ///< claim cursor of each job, one per 256-byte block
};

struct BatchArgs
{
    const JobDesc* jobs    = nullptr;
    BatchCtl*      ctl     = nullptr;
    std::uint32_t  njobs   = 0;
};

/// @brief Claim a guided chunk of job @p j: returns the count.
LOOM_DEVICE std::uint32_t
claim_job_chunk(BatchCtl* ctl, const JobDesc* J, std::uint32_t j, std::uint32_t chunk,
                std::uint64_t& base)
{
    std::uint64_t b = 0;
    std::uint32_t got = 0;
    if(lane_id() == 0)
    {
        const std::uint64_t e = J[j].end;
        const std::uint64_t k = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->next[j].v),
                                          static_cast<unsigned long long>(chunk));
        if(k < e)
        {
            b   = k;
            got = static_cast<std::uint32_t>(e - k < chunk ? e - k : chunk);
        }
    }
    base = uniform64(b);
    return uniform(got);
}

/// @brief Run job @p j (kernel K) on this wave until the job has no more
/// threads to admit and every lane halted.
template <typename K, int P>
LOOM_DEVICE void
run_job(const BatchArgs& a, std::uint32_t j, const CtlArgs& c, WaveCtx& w, std::uint64_t& admitted)
{
    const JobDesc*      J   = a.jobs;
    const std::uint64_t len = J[j].end - J[j].begin;
    // chunk: the job spread over about four claims per wave that starts on
    // it (gridDim.x / njobs waves do), one batch at least, 512 at most
    const std::uint64_t per = len * a.njobs / (4ull * gridDim.x);
    const std::uint32_t chunk = static_cast<std::uint32_t>(per < 64 ? 64 : (per > 512 ? 512 : per));
    std::uint64_t rbase = 0;
    std::uint32_t rcount = 0;
    bool          exhausted = false;

    Regs<K>       r;
    std::uint32_t pc  = HALT;
    std::uint64_t tid = 0;
    bool          has = false;
    bool          force = true;

    auto take = [&](std::uint32_t want, std::uint64_t& base) -> std::uint32_t
    {
        if(rcount == 0 && !exhausted)
        {
            rcount = claim_job_chunk(a.ctl, J, j, chunk, rbase);
            if(!rcount) exhausted = true;
        }
        const std::uint32_t g = rcount < want ? rcount : want;
        base                  = rbase;
        rbase += g;
        rcount -= g;
        return g;
    };

    for(;;)
    {
        flush_if_high(w);
        const std::uint64_t live = ballot(has);
        if(~live)
        {
            std::uint64_t       base = 0;
            const std::uint64_t empty = ~live;
            const std::uint32_t got   = take(popc(empty), base);
            if(got)
            {
                const std::uint32_t k = rank_in(empty);
                if(!has && k < got)
                {
                    tid = base + k;
                    load_inputs<K>(r, J[j].in, tid);
                    pc  = 0;
                    has = true;
                }
                admitted += got;
                force = true;
            }
        }
        if(!ballot(has)) return;
        run_policy<K, P>(r, pc, 0, exhausted ? 0u : c.live_thr, force, w);
        force = false;
        if(has && pc == HALT)
        {
            store_outputs<K>(r, J[j].out, tid);
            has = false;
        }
        if(w.over) return;
    }
}

template <typename... Ks>
struct BatchSet
{
    static constexpr std::uint32_t size = sizeof...(Ks);
};

/// @brief The persistent batch kernel over the kernels @p Ks (policy @p P).
template <int P, typename... Ks>
__global__ __launch_bounds__(64) void
batch_kernel(CtlArgs c, BatchArgs a)
{
    (sct_table_init<Ks>(), ...);
    WaveCtx w(c);
    std::uint64_t admitted = 0;
    const std::uint32_t nj = a.njobs;
    // each wave walks the job list from its own start, so jobs of every size
    // are worked on from the beginning
    for(std::uint32_t step = 0; step < nj && !w.over; ++step)
    {
        const std::uint32_t j = (blockIdx.x + step) % nj;
        // skip exhausted jobs cheaply
        if(uniform64(__hip_atomic_load(&a.ctl->next[j].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >=
           a.jobs[j].end)
            continue;
        const std::uint32_t m = a.jobs[j].member;
        std::uint32_t       idx = 0;
        (
            [&]
            {
                if(m == idx) run_job<Ks, P>(a, j, c, w, admitted);
                ++idx;
            }(),
            ...);
    }
    flush_stats(w);
    if(c.stats && lane_id() == 0)
        atomicAdd(&stat_slot(c)[ST_ADMITTED], static_cast<unsigned long long>(admitted));
    raise_errors(w.err, w.over, c);
}

/// @brief Reset the claim cursors of a batch (one thread per job).
inline __global__ void
batch_reset_kernel(BatchArgs a)
{
    const std::uint32_t j = threadIdx.x;
    if(j < a.njobs) a.ctl->next[j].v = a.jobs[j].begin;
}
} // namespace loom::hipvm
