// -*- C++ -*-
/// @file runtime.cpp
/// @brief The C ABI (include/loom_amd.h): device set-up, kernel registry,
/// device memory, execution contexts, and the compact / hybrid / batched run
/// drivers.
///
/// Work is issued on execution contexts (loom_context): each owns a kernel
/// stream and a copy stream (default priority), its error flags and
/// counters, and grow-only workspaces, so repeated runs do not allocate.
/// A context's lock is held for the whole of each call on it, so one
/// context may be shared by several host threads (their calls serialise),
/// and host threads with contexts of their own run concurrently.  The entry
/// points without a context argument use the default context that
/// loom_init creates.
#include "runtime.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

namespace loom::hipvm
{
namespace
{
std::vector<KernelEntry>&
registry()
{
    static std::vector<KernelEntry> r;
    return r;
}

std::vector<BatchEntry>&
batch_registry()
{
    static std::vector<BatchEntry> r;
    return r;
}

thread_local std::string g_error;

/// Device memory of loom_run_noalloc: the hybrid control block at the start,
/// the two regroup queue sides after it.  One copy per GPU, held until the
/// process ends; host code uses g_scratch (its device address, looked up
/// at loom_init), never the array's name.
constexpr std::size_t kScratchBytes = std::size_t(16) << 20;
constexpr std::size_t kScratchCtl   = (sizeof(HybridCtl) + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void*                                           g_scratch = nullptr;

Device        g_dev;
std::mutex    g_init_mutex;
loom_context* g_default = nullptr;

IoArgs
make_io(const KernelEntry& e, const void* const* in, void* const* out)
{
    IoArgs io{};
    for(int k = 0; k < e.info.n_inputs; ++k) io.in[k] = in[k];
    for(int k = 0; k < e.info.n_outputs; ++k) io.out[k] = out[k];
    return io;
}

int
validate_run(const KernelEntry& e, int mode, const void* const* inputs, void* const* outputs,
             std::uint64_t n, const loom_params* params)
{
    if(mode != LOOM_MODE_COMPACT && mode != LOOM_MODE_HYBRID)
        return fail(LOOM_E_ARG, "unknown mode %d", mode);
    if(n > 0xFFFFFFFFull) return fail(LOOM_E_CAPACITY, "at most 2^32-1 threads per run");
    if(n && (!inputs || !outputs)) return fail(LOOM_E_ARG, "null column array");
    for(int k = 0; n && k < e.info.n_inputs; ++k)
        if(!inputs[k]) return fail(LOOM_E_ARG, "input column %d is null", k);
    for(int k = 0; n && k < e.info.n_outputs; ++k)
        if(!outputs[k]) return fail(LOOM_E_ARG, "output column %d is null", k);
    return validate_params(params);
}

/// @brief Resolve a context argument: NULL means the default context.
int
context_of(loom_context* ctx, loom_context*& out)
{
    if(int r = ensure_device()) return r;
    out = ctx ? ctx : g_default;
    if(!out) return fail(LOOM_E_NODEVICE, "no default context");
    return 0;
}

int
context_init(loom_context& x)
{
    LOOM_HIP(hipSetDevice(g_dev.id));
    LOOM_HIP(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
    LOOM_HIP(hipStreamCreateWithFlags(&x.copy, hipStreamNonBlocking));
    LOOM_HIP(hipMalloc(&x.err, sizeof(std::uint32_t)));
    LOOM_HIP(hipMalloc(&x.stats, sizeof(std::uint64_t) * kStatWords));
    LOOM_HIP(hipHostMalloc(&x.host_err, sizeof(std::uint32_t), hipHostMallocDefault));
    LOOM_HIP(hipHostMalloc(&x.host_stats, sizeof(std::uint64_t) * kStatWords, hipHostMallocDefault));
    LOOM_HIP(hipHostMalloc(&x.host_ctl, sizeof(HybridCtl), hipHostMallocDefault));
    for(auto& ev: x.ev) LOOM_HIP(hipEventCreate(&ev));
    return 0;
}

/// @brief Release everything a context holds (after its work finished).
void
context_release(loom_context& x)
{
    if(x.stream) (void)hipStreamSynchronize(x.stream);
    if(x.copy) (void)hipStreamSynchronize(x.copy);
    for(auto& w: x.ws)
    {
        if(w.ctl) (void)hipFree(w.ctl);
        for(auto q: w.q)
            if(q) (void)hipFree(q);
    }
    for(void* p: x.hostcol_ptr)
        if(p) (void)hipFree(p);
    if(x.memregs) (void)hipFree(x.memregs);
    if(x.wgr_next) (void)hipFree(x.wgr_next);
    if(x.jobs) (void)hipFree(x.jobs);
    if(x.bctl) (void)hipFree(x.bctl);
    for(auto& t: x.tickets) (void)hipEventDestroy(t.second);
    for(auto e: x.spare_events) (void)hipEventDestroy(e);
    for(auto e: x.pass_ev) (void)hipEventDestroy(e);
    for(auto e: x.ev)
        if(e) (void)hipEventDestroy(e);
    if(x.err) (void)hipFree(x.err);
    if(x.stats) (void)hipFree(x.stats);
    if(x.host_err) (void)hipHostFree(x.host_err);
    if(x.host_stats) (void)hipHostFree(x.host_stats);
    if(x.host_ctl) (void)hipHostFree(x.host_ctl);
    if(x.stream) (void)hipStreamDestroy(x.stream);
    if(x.copy) (void)hipStreamDestroy(x.copy);
}

int
finish_stats(loom_context& x, loom_run_stats* st, const loom_params* p)
{
    LOOM_HIP(hipMemcpyAsync(x.host_err, x.err, sizeof(std::uint32_t), hipMemcpyDeviceToHost,
                            x.stream));
    if(p && p->collect_stats)
        LOOM_HIP(hipMemcpyAsync(x.host_stats, x.stats, sizeof(std::uint64_t) * kStatWords,
                                hipMemcpyDeviceToHost, x.stream));
    LOOM_HIP(hipStreamSynchronize(x.stream));
    if(st)
    {
        st->errors = *x.host_err;
        if(p && p->collect_stats) sum_counters(x.host_stats, st->counters);
    }
    if(*x.host_err) return fail(LOOM_E_RUNTIME, "kernel raised error flags 0x%x", *x.host_err);
    return 0;
}

int
run_compact(loom_context& x, const KernelEntry& e, const IoArgs& io, CtlArgs c, int policy,
            bool memory, loom_run_stats* st)
{
    MemArgs m;
    if(memory)
    {
        // one row of register cells per thread, grown on demand
        const std::size_t need = std::max<std::size_t>(
            1, static_cast<std::size_t>(c.n) * std::max<std::uint32_t>(e.state_bytes, 4));
        if(need > x.memregs_bytes)
        {
            LOOM_HIP(hipStreamSynchronize(x.stream));
            if(x.memregs) LOOM_HIP(hipFree(x.memregs));
            x.memregs       = nullptr;
            x.memregs_bytes = 0;
            LOOM_HIP(hipMalloc(&x.memregs, need));
            x.memregs_bytes = need;
        }
        m.base = x.memregs;
        m.rows = c.n;
    }
    LOOM_HIP(hipEventRecord(x.ev[0], x.stream));
    if(memory)
        LOOM_HIP(e.compact_mem(io, c, m, x.stream));
    else
        LOOM_HIP(e.compact[policy](io, c, x.stream));
    LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
    LOOM_HIP(hipEventSynchronize(x.ev[1]));
    float ms = 0;
    LOOM_HIP(hipEventElapsedTime(&ms, x.ev[0], x.ev[1]));
    if(st)
    {
        st->ms_total      = ms;
        st->ms_main       = ms;
        st->main_launches = 1;
        st->passes        = 1;
        st->grid_waves    = static_cast<std::uint32_t>((c.n + 63) / 64);
    }
    return 0;
}

/// @brief The persistent grid of a hybrid run: resident 64-lane workgroups
/// (waves_per_cu per CU, capped by occupancy and by the work).
unsigned
hybrid_grid(const KernelEntry& e, int policy, bool regroup, const loom_params* p, std::uint64_t n)
{
    const int occ = e.hybrid_occupancy[policy](regroup);
    int       wpc = p && p->waves_per_cu ? static_cast<int>(p->waves_per_cu) : occ;
    wpc           = std::max(1, std::min(wpc, occ));
    unsigned grid = static_cast<unsigned>(g_dev.cus * wpc);
    const std::uint64_t need_waves = (n + 63) / 64;
    if(need_waves < grid) grid = static_cast<unsigned>(std::max<std::uint64_t>(1, need_waves));
    return grid;
}

/// @brief Queue capacity (rows per label, a multiple of 64) of a hybrid run.
std::uint64_t
hybrid_qcap(const loom_params* p, unsigned grid, std::uint64_t n)
{
    std::uint64_t qcap = p && p->queue_capacity ? p->queue_capacity :
                                                  std::uint64_t(grid) * 64ull * 4ull;
    qcap = std::min<std::uint64_t>(qcap, std::max<std::uint64_t>(n, 64));
    return (qcap + 63) / 64 * 64;
}

/// @brief Launch sequence of a hybrid run over the workspace @p h (control
/// block and queue sides) on @p x's stream.
int
run_hybrid_on(loom_context& x, const KernelEntry& e, const IoArgs& io, CtlArgs c, int policy,
              unsigned grid, const HybridArgs& h, loom_run_stats* st, bool fixed_events = false)
{
    const std::uint32_t nb = static_cast<std::uint32_t>(e.info.n_blocks);
    LOOM_HIP(hipEventRecord(x.ev[0], x.stream));
    hipLaunchKernelGGL(hybrid_setup_kernel, dim3(1), dim3(64), 0, x.stream, h.ctl, nb, h.qcap,
                       c.n, 1u);
    LOOM_HIP(hipGetLastError());

    std::uint32_t passes = 0, check_every = 1;
    constexpr std::uint32_t max_passes = 1u << 20;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    std::size_t ev_used = 0;
    auto next_ev = [&]() -> hipEvent_t
    {
        // without event creation: one timed pass on the context's own events
        if(fixed_events) return ev_used < 2 ? x.ev[2 + ev_used++] : nullptr;
        if(ev_used == x.pass_ev.size())
        {
            hipEvent_t ev{};
            if(hipEventCreate(&ev) != hipSuccess) return nullptr;
            x.pass_ev.push_back(ev);
        }
        return x.pass_ev[ev_used++];
    };
    if(!c.regroup)
    {
        // One pass does all the work: every wave takes input until none is
        // left and finishes its remaining lanes in place (a wave over its
        // step budget raises the error flag instead), so no control-block
        // read-back and host round trip is needed to learn that it is done.
        hipEvent_t a = next_ev(), b = next_ev();
        if(!a || !b) return fail(LOOM_E_RUNTIME, "event creation failed");
        LOOM_HIP(hipEventRecord(a, x.stream));
        LOOM_HIP(e.hybrid_pass[policy](io, c, h, grid, x.stream));
        LOOM_HIP(hipEventRecord(b, x.stream));
        evs.emplace_back(a, b);
        passes = 1;
        LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
    }
    else
        for(;;)
        {
            for(std::uint32_t k = 0; k < check_every; ++k)
            {
                if(fixed_events)
                    LOOM_HIP(e.hybrid_pass[policy](io, c, h, grid, x.stream));
                else
                {
                    hipEvent_t a = next_ev(), b = next_ev();
                    if(!a || !b) return fail(LOOM_E_RUNTIME, "event creation failed");
                    LOOM_HIP(hipEventRecord(a, x.stream));
                    LOOM_HIP(e.hybrid_pass[policy](io, c, h, grid, x.stream));
                    LOOM_HIP(hipEventRecord(b, x.stream));
                    evs.emplace_back(a, b);
                }
                hipLaunchKernelGGL(hybrid_setup_kernel, dim3(1), dim3(64), 0, x.stream, h.ctl, nb,
                                   h.qcap, c.n, 0u);
                LOOM_HIP(hipGetLastError());
                ++passes;
            }
            LOOM_HIP(hipMemcpyAsync(x.host_ctl, h.ctl, sizeof(HybridCtl), hipMemcpyDeviceToHost,
                                    x.stream));
            // the end of the run, if this check finds it done: recorded
            // before the host waits, so the host's round trip is not timed
            LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
            LOOM_HIP(hipStreamSynchronize(x.stream));
            if(x.host_ctl->done) break;
            if(passes >= max_passes)
                return fail(LOOM_E_RUNTIME, "hybrid run exceeded %u passes", max_passes);
            check_every = std::min<std::uint32_t>(check_every * 2, 8);
        }
    LOOM_HIP(hipEventSynchronize(x.ev[1]));
    float total = 0;
    LOOM_HIP(hipEventElapsedTime(&total, x.ev[0], x.ev[1]));
    double main_ms = 0;
    for(auto& [a, b]: evs)
    {
        float ms = 0;
        LOOM_HIP(hipEventElapsedTime(&ms, a, b));
        main_ms += ms;
    }
    // passes timed only as a whole (fixed events, regrouping run)
    if(evs.empty()) main_ms = total;
    if(st)
    {
        st->ms_total      = total;
        st->ms_main       = main_ms;
        st->main_launches = evs.empty() ? passes : static_cast<std::uint32_t>(evs.size());
        st->passes        = passes;
        st->grid_waves    = grid;
    }
    return 0;
}

/// @brief A run of the workgroup-regrouping kernel: a persistent grid of
/// resident workgroups, one launch.
int
run_wgr(loom_context& x, const KernelEntry& e, const IoArgs& io, CtlArgs c, const loom_params* p,
        loom_run_stats* st)
{
    const int which = p->reconv == POLICY_SUCCESSOR_MAJOR ? 0 : 1;
    const int waves = p->regroup_waves == 8 ? 8 : kWgrWaves;
    if(!x.wgr_next) LOOM_HIP(hipMalloc(&x.wgr_next, sizeof(std::uint64_t)));
    LOOM_HIP(hipMemsetAsync(x.wgr_next, 0, sizeof(std::uint64_t), x.stream));
    WgrArgs g;
    g.next  = x.wgr_next;
    g.phase = p->regroup_phase ? p->regroup_phase : 64;
    const std::uint64_t lanes = 64ull * static_cast<std::uint64_t>(waves);
    const std::uint64_t need  = (c.n + lanes - 1) / lanes;
    unsigned grid = static_cast<unsigned>(g_dev.cus * e.wgr_occupancy[which](waves));
    if(need < grid) grid = static_cast<unsigned>(std::max<std::uint64_t>(1, need));
    LOOM_HIP(hipEventRecord(x.ev[0], x.stream));
    LOOM_HIP(e.wgr[which](io, c, g, waves, grid, x.stream));
    LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
    LOOM_HIP(hipEventSynchronize(x.ev[1]));
    float ms = 0;
    LOOM_HIP(hipEventElapsedTime(&ms, x.ev[0], x.ev[1]));
    if(st)
    {
        st->ms_total      = ms;
        st->ms_main       = ms;
        st->main_launches = 1;
        st->passes        = 1;
        st->grid_waves    = grid * static_cast<std::uint32_t>(waves);
    }
    return 0;
}

int
run_hybrid(loom_context& x, int kidx, const KernelEntry& e, const IoArgs& io, CtlArgs c,
           const loom_params* p, loom_run_stats* st)
{
    if(p && p->regroup == LOOM_REGROUP_WORKGROUP) return run_wgr(x, e, io, c, p, st);
    const int      policy = p ? static_cast<int>(p->reconv) : static_cast<int>(POLICY_SWEEP);
    const unsigned grid   = hybrid_grid(e, policy, c.regroup != 0, p, c.n);
    const std::uint64_t qcap = hybrid_qcap(p, grid, c.n);
    if(qcap > 0xFFFFFFC0ull) return fail(LOOM_E_CAPACITY, "queue capacity too large");

    if(x.ws.size() < registry().size()) x.ws.resize(registry().size());
    auto&             ws   = x.ws[kidx];
    const std::size_t side = std::size_t(qcap) * e.info.queue_row_bytes;
    if(!ws.ctl) LOOM_HIP(hipMalloc(&ws.ctl, sizeof(HybridCtl)));
    // queues are read only by runs that regroup
    if(c.regroup && ws.bytes < side)
    {
        for(auto& q: ws.q)
            if(q)
            {
                LOOM_HIP(hipFree(q));
                q = nullptr;
            }
        ws.bytes = 0;
        for(auto& q: ws.q) LOOM_HIP(hipMalloc(&q, side));
        ws.bytes = side;
    }
    HybridArgs h;
    h.ctl  = ws.ctl;
    h.q[0] = ws.q[0];
    h.q[1] = ws.q[1];
    h.qcap = qcap;
    return run_hybrid_on(x, e, io, c, policy, grid, h, st);
}

/// @brief A hybrid run whose control block and queues are the scratch
/// array (no allocation): fails with LOOM_E_CAPACITY when the queues of a
/// regrouping run do not fit.
int
run_hybrid_scratch(loom_context& x, const KernelEntry& e, const IoArgs& io, CtlArgs c,
                   const loom_params* p, loom_run_stats* st)
{
    if(!g_scratch) return fail(LOOM_E_NODEVICE, "scratch memory not set up");
    const int      policy = p ? static_cast<int>(p->reconv) : static_cast<int>(POLICY_SWEEP);
    const unsigned grid   = hybrid_grid(e, policy, c.regroup != 0, p, c.n);
    std::uint64_t  qcap   = hybrid_qcap(p, grid, c.n);
    char*          base   = static_cast<char*>(g_scratch);
    HybridArgs     h;
    h.ctl  = reinterpret_cast<HybridCtl*>(base);
    h.qcap = qcap;
    if(c.regroup)
    {
        const std::uint64_t side = qcap * e.info.queue_row_bytes;
        if(2 * side > kScratchBytes - kScratchCtl)
            return fail(LOOM_E_CAPACITY,
                        "regroup queues need %llu bytes, the scratch array holds %zu",
                        static_cast<unsigned long long>(2 * side), kScratchBytes - kScratchCtl);
        h.q[0] = base + kScratchCtl;
        h.q[1] = base + kScratchCtl + side;
    }
    return run_hybrid_on(x, e, io, c, policy, grid, h, st, true);
}

int
run_on(loom_context& x, int kernel, int mode, const void* const* inputs, void* const* outputs,
       std::uint64_t n, const loom_params* params, loom_run_stats* stats, bool scratch = false)
{
    if(int r = check_kernel(kernel)) return r;
    const KernelEntry& e = registry()[kernel];
    if(int r = validate_run(e, mode, inputs, outputs, n, params)) return r;
    if(stats) *stats = loom_run_stats{};
    const IoArgs  io = n ? make_io(e, inputs, outputs) : IoArgs{};
    const CtlArgs c  = make_ctl(params, n, mode, x.err, x.stats);
    LOOM_HIP(hipMemsetAsync(x.err, 0, sizeof(std::uint32_t), x.stream));
    if(c.stats) LOOM_HIP(hipMemsetAsync(x.stats, 0, sizeof(std::uint64_t) * kStatWords, x.stream));
    int r = 0;
    const bool memory = params && params->register_file == LOOM_REGFILE_MEMORY;
    if(memory && mode != LOOM_MODE_COMPACT)
        return fail(LOOM_E_ARG, "a register file in memory runs in the compact form");
    if(memory && scratch)
        return fail(LOOM_E_ARG, "a register file in memory needs an allocation; use loom_run");
    if(params && params->regroup == LOOM_REGROUP_WORKGROUP && scratch)
        return fail(LOOM_E_ARG, "workgroup regrouping needs an allocation; use loom_run");
    if(mode == LOOM_MODE_COMPACT)
        r = run_compact(x, e, io, c, params ? static_cast<int>(params->reconv) : 0, memory, stats);
    else if(scratch)
        r = run_hybrid_scratch(x, e, io, c, params, stats);
    else
        r = run_hybrid(x, kernel, e, io, c, params, stats);
    if(r) return r;
    return finish_stats(x, stats, params);
}

int
run_host_on(loom_context& x, int kernel, int mode, const void* const* inputs,
            void* const* outputs, std::uint64_t n, const loom_params* params,
            loom_run_stats* stats)
{
    if(int r = check_kernel(kernel)) return r;
    const KernelEntry& e = registry()[kernel];
    // validate before allocating or copying anything
    if(int r = validate_run(e, mode, inputs, outputs, n, params)) return r;
    const int          ni = e.info.n_inputs, no = e.info.n_outputs;
    std::vector<void*> din(ni), dout(no);
    if(x.hostcol_ptr.size() < static_cast<std::size_t>(ni + no))
    {
        x.hostcol_ptr.resize(ni + no, nullptr);
        x.hostcol_bytes.resize(ni + no, 0);
    }
    for(int k = 0; k < ni + no; ++k)
    {
        const int         t     = k < ni ? e.info.input_type[k] : e.info.output_type[k - ni];
        const std::size_t bytes = std::max<std::size_t>(1, n * type_size(t));
        if(x.hostcol_bytes[k] < bytes)
        {
            if(x.hostcol_ptr[k]) LOOM_HIP(hipFree(x.hostcol_ptr[k]));
            x.hostcol_ptr[k]   = nullptr;
            x.hostcol_bytes[k] = 0;
            LOOM_HIP(hipMalloc(&x.hostcol_ptr[k], bytes));
            x.hostcol_bytes[k] = bytes;
        }
        (k < ni ? din[k] : dout[k - ni]) = x.hostcol_ptr[k];
    }
    for(int k = 0; k < ni && n; ++k)
        LOOM_HIP(hipMemcpyAsync(din[k], inputs[k], n * type_size(e.info.input_type[k]),
                                hipMemcpyHostToDevice, x.stream));
    if(int r = run_on(x, kernel, mode, din.data(), dout.data(), n, params, stats)) return r;
    for(int k = 0; k < no && n; ++k)
        LOOM_HIP(hipMemcpyAsync(outputs[k], dout[k], n * type_size(e.info.output_type[k]),
                                hipMemcpyDeviceToHost, x.stream));
    LOOM_HIP(hipStreamSynchronize(x.stream));
    return 0;
}

/// @brief Broadcast one element over a column (grid-stride, vector stores).
template <typename T>
__global__ void
fill_kernel(T* dst, T value, std::uint64_t count)
{
    const std::uint64_t stride = std::uint64_t(gridDim.x) * blockDim.x;
    for(std::uint64_t i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = value;
}

template <typename T>
hipError_t
launch_fill(void* dst, const void* pattern, std::uint64_t count, hipStream_t s)
{
    T v;
    std::memcpy(&v, pattern, sizeof(T));
    const std::uint64_t blocks = std::min<std::uint64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<T*>(dst), v, count);
    return hipGetLastError();
}

__global__ void
zero_counters_kernel(std::uint32_t* err, std::uint64_t* stats, std::uint32_t nstats)
{
    if(threadIdx.x == 0) *err = 0;
    if(stats)
        for(std::uint32_t i = threadIdx.x; i < nstats; i += blockDim.x) stats[i] = 0;
}

int
batch_jobs(int set, const loom_job* jobs, int njobs, std::vector<JobDesc>& out)
{
    if(set < 0 || set >= static_cast<int>(batch_registry().size()))
        return fail(LOOM_E_KERNEL, "unknown batch set %d", set);
    if(njobs < 1 || njobs > static_cast<int>(kMaxJobs) || !jobs)
        return fail(LOOM_E_ARG, "a batch holds 1..%u jobs", kMaxJobs);
    const BatchEntry& b = batch_registry()[set];
    out.assign(njobs, JobDesc{});
    for(int i = 0; i < njobs; ++i)
    {
        const loom_job& j = jobs[i];
        if(int r = check_kernel(j.kernel)) return r;
        const KernelEntry& e = registry()[j.kernel];
        std::uint32_t      m = b.size;
        for(std::uint32_t k = 0; k < b.size; ++k)
            if(b.hash[k] == e.info.hash) m = k;
        if(m == b.size)
            return fail(LOOM_E_KERNEL, "kernel %s is not in batch set %s", e.info.name, b.name);
        if(j.end < j.begin || j.end > 0xFFFFFFFFull)
            return fail(LOOM_E_ARG, "job %d: bad thread range", i);
        out[i].member = m;
        out[i].begin  = j.begin;
        out[i].end    = j.end;
        for(int k = 0; k < e.info.n_inputs; ++k)
        {
            if(!j.inputs[k] && j.end > j.begin) return fail(LOOM_E_ARG, "job %d input %d null", i, k);
            out[i].in[k] = j.inputs[k];
        }
        for(int k = 0; k < e.info.n_outputs; ++k)
        {
            if(!j.outputs[k] && j.end > j.begin) return fail(LOOM_E_ARG, "job %d output %d null", i, k);
            out[i].out[k] = j.outputs[k];
        }
    }
    return 0;
}

int
batch_policy(const loom_params* p)
{
    return p ? static_cast<int>(p->reconv) : static_cast<int>(POLICY_SWEEP);
}

unsigned
batch_grid(const BatchEntry& b, const loom_params* p, std::uint64_t threads)
{
    const int occ = b.occupancy[batch_policy(p)]();
    int       wpc = p && p->waves_per_cu ? static_cast<int>(p->waves_per_cu) : occ;
    wpc           = std::max(1, std::min(wpc, occ));
    unsigned grid = static_cast<unsigned>(g_dev.cus * wpc);
    const std::uint64_t need = (threads + 63) / 64;
    if(need < grid) grid = static_cast<unsigned>(std::max<std::uint64_t>(1, need));
    return grid;
}

int
batch_run_on(loom_context& x, int set, const loom_job* jobs, int njobs, const loom_params* params,
             loom_run_stats* stats)
{
    if(int r = validate_params(params)) return r;
    if(params && params->register_file != LOOM_REGFILE_VGPR)
        return fail(LOOM_E_ARG, "batched launches keep registers in VGPRs");
    if(params && params->regroup == LOOM_REGROUP_WORKGROUP)
        return fail(LOOM_E_ARG, "batched launches do not regroup within workgroups");
    std::vector<JobDesc> jd;
    if(int r = batch_jobs(set, jobs, njobs, jd)) return r;
    if(stats) *stats = loom_run_stats{};
    if(!x.jobs) LOOM_HIP(hipMalloc(&x.jobs, sizeof(JobDesc) * kMaxJobs));
    if(!x.bctl) LOOM_HIP(hipMalloc(&x.bctl, sizeof(BatchCtl)));
    LOOM_HIP(hipMemcpyAsync(x.jobs, jd.data(), sizeof(JobDesc) * jd.size(), hipMemcpyHostToDevice,
                            x.stream));
    std::uint64_t threads = 0;
    for(const auto& j: jd) threads += j.end - j.begin;
    const BatchEntry& b = batch_registry()[set];
    const CtlArgs     c = make_ctl(params, threads, LOOM_MODE_HYBRID, x.err, x.stats);
    BatchArgs         a;
    a.jobs  = x.jobs;
    a.ctl   = x.bctl;
    a.njobs = static_cast<std::uint32_t>(njobs);
    const unsigned grid = batch_grid(b, params, threads);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, x.stream, x.err, x.stats,
                       static_cast<std::uint32_t>(kStatWords));
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(kMaxJobs), 0, x.stream, a);
    LOOM_HIP(hipGetLastError());
    LOOM_HIP(hipEventRecord(x.ev[0], x.stream));
    LOOM_HIP(b.launch[batch_policy(params)](c, a, grid, x.stream));
    LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
    LOOM_HIP(hipEventSynchronize(x.ev[1]));
    float ms = 0;
    LOOM_HIP(hipEventElapsedTime(&ms, x.ev[0], x.ev[1]));
    if(stats)
    {
        stats->ms_total      = ms;
        stats->ms_main       = ms;
        stats->main_launches = 1;
        stats->passes        = 1;
        stats->grid_waves    = grid;
    }
    return finish_stats(x, stats, params);
}

int
copy_async_on(loom_context& x, void* dst, const void* src, size_t bytes, int kind,
              std::uint64_t* ticket)
{
    if(!ticket) return fail(LOOM_E_ARG, "null ticket");
    if(bytes && (!dst || !src)) return fail(LOOM_E_ARG, "null pointer");
    if(kind != LOOM_COPY_H2D && kind != LOOM_COPY_D2H)
        return fail(LOOM_E_ARG, "unknown copy kind %d", kind);
    if(bytes)
        LOOM_HIP(hipMemcpyAsync(dst, src, bytes,
                                kind == LOOM_COPY_H2D ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                                x.copy));
    hipEvent_t ev{};
    if(!x.spare_events.empty())
    {
        ev = x.spare_events.back();
        x.spare_events.pop_back();
    }
    else
        LOOM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    LOOM_HIP(hipEventRecord(ev, x.copy));
    *ticket = ++x.last_ticket;
    x.tickets.emplace_back(*ticket, ev);
    return 0;
}

int
copy_status_on(loom_context& x, std::uint64_t ticket, int* done)
{
    if(!done) return fail(LOOM_E_ARG, "null result");
    while(!x.tickets.empty())
    {
        const hipError_t q = hipEventQuery(x.tickets.front().second);
        if(q == hipErrorNotReady) break;
        if(q != hipSuccess) return fail(static_cast<int>(q), "copy: %s", hipGetErrorString(q));
        x.spare_events.push_back(x.tickets.front().second);
        x.tickets.pop_front();
    }
    *done = x.tickets.empty() || x.tickets.front().first > ticket;
    return 0;
}
} // namespace

const std::vector<KernelEntry>&
kernels()
{
    return registry();
}

const std::vector<BatchEntry>&
batch_sets()
{
    return batch_registry();
}

const Device&
device()
{
    return g_dev;
}

loom_context*
default_context()
{
    return g_default;
}

int
fail(int code, const char* fmt, ...)
{
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int
ensure_device()
{
    if(!g_dev.ready) return fail(LOOM_E_NODEVICE, "loom_init has not been called");
    return 0;
}

int
check_kernel(int k)
{
    if(k < 0 || k >= static_cast<int>(registry().size()))
        return fail(LOOM_E_KERNEL, "unknown kernel index %d", k);
    return 0;
}

int
type_size(int t)
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

int
validate_params(const loom_params* params)
{
    if(params && (params->threshold > 63 || params->live_threshold > 63))
        return fail(LOOM_E_ARG, "thresholds must be < 64");
    if(params && params->reconv >= static_cast<std::uint32_t>(kNumPolicies))
        return fail(LOOM_E_ARG, "unknown reconvergence policy");
    if(params && params->register_file > LOOM_REGFILE_MEMORY)
        return fail(LOOM_E_ARG, "unknown register file %u", params->register_file);
    if(params && params->register_file == LOOM_REGFILE_MEMORY && params->reconv != POLICY_SWEEP)
        return fail(LOOM_E_ARG, "a register file in memory runs under the sweep policy (reconv 2)");
    if(params && params->regroup > LOOM_REGROUP_WORKGROUP)
        return fail(LOOM_E_ARG, "unknown regroup mode %u", params->regroup);
    if(params && params->regroup == LOOM_REGROUP_WORKGROUP && params->reconv == POLICY_TOPO)
        return fail(LOOM_E_ARG, "workgroup regrouping runs under reconv 0 or 2");
    if(params && params->regroup_waves != 0 && params->regroup_waves != 8 &&
       params->regroup_waves != 16)
        return fail(LOOM_E_ARG, "a regrouping workgroup has 16 or 8 waves, not %u",
                    params->regroup_waves);
    return 0;
}

CtlArgs
make_ctl(const loom_params* p, std::uint64_t n, int mode, std::uint32_t* err, std::uint64_t* stats)
{
    CtlArgs c;
    c.n         = n;
    c.max_steps = p && p->max_steps ? p->max_steps : (1ull << 40);
    c.thr       = p ? p->threshold : 0;
    c.live_thr  = p ? p->live_threshold : (mode == LOOM_MODE_HYBRID ? 48 : 0);
    c.refill    = p ? p->refill : 1;
    c.regroup   = p ? p->regroup : 0;
    c.err       = err;
    c.stats     = p && p->collect_stats ? stats : nullptr;
    return c;
}

void
sum_counters(const std::uint64_t* host_stats, std::uint64_t* out)
{
    for(std::size_t i = 0; i < kNumStats; ++i)
    {
        std::uint64_t sum = 0;
        for(std::size_t k = 0; k < kStatSlots; ++k) sum += host_stats[k * kNumStats + i];
        out[i] = sum;
    }
}

bool
register_kernel(const KernelEntry& e)
{
    registry().push_back(e);
    return true;
}

bool
register_batch(const BatchEntry& e)
{
    batch_registry().push_back(e);
    return true;
}
} // namespace loom::hipvm

using namespace loom::hipvm;

namespace
{
/// @brief Holds a context's lock for the rest of a call.
struct Locked
{
    loom_context*                x = nullptr;
    std::unique_lock<std::mutex> lk;
};

/// @brief Resolve @p ctx (NULL: the default context) and lock it.
int
lock_context(loom_context* ctx, Locked& l)
{
    if(int r = context_of(ctx, l.x)) return r;
    l.lk = std::unique_lock<std::mutex>(l.x->mu);
    return 0;
}
} // namespace

extern "C" {

const char*
loom_last_error(void)
{
    return g_error.c_str();
}

int
loom_init(int device)
{
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if(g_dev.ready) return device == g_dev.id ? 0 : fail(LOOM_E_ARG, "already initialised");
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(LOOM_E_NODEVICE, "no HIP device");
    if(device < 0 || device >= count) return fail(LOOM_E_ARG, "device %d out of range", device);
    LOOM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop{};
    LOOM_HIP(hipGetDeviceProperties(&prop, device));
    g_dev.id  = device;
    g_dev.cus = prop.multiProcessorCount;
    // the device address of this GPU's scratch array (host code must not
    // pass the array's own name to a kernel)
    LOOM_HIP(hipGetSymbolAddress(&g_scratch, HIP_SYMBOL(scratch)));
    auto x    = std::make_unique<loom_context>();
    if(int r = context_init(*x))
    {
        context_release(*x);
        return r;
    }
    g_default   = x.release();
    g_dev.ready = true;
    return 0;
}

int
loom_context_create(loom_context** ctx)
{
    if(int r = ensure_device()) return r;
    if(!ctx) return fail(LOOM_E_ARG, "null context");
    auto x = std::make_unique<loom_context>();
    if(int r = context_init(*x))
    {
        context_release(*x);
        return r;
    }
    *ctx = x.release();
    return 0;
}

int
loom_context_destroy(loom_context* ctx)
{
    if(!ctx) return 0;
    if(ctx == g_default) return fail(LOOM_E_ARG, "the default context is not destroyed");
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        context_release(*ctx);
    }
    delete ctx;
    return 0;
}

int
loom_device_info(int* compute_units, int* wave_size, size_t* global_mem, char* name,
                 size_t name_len)
{
    if(int r = ensure_device()) return r;
    hipDeviceProp_t prop{};
    LOOM_HIP(hipGetDeviceProperties(&prop, g_dev.id));
    if(compute_units) *compute_units = prop.multiProcessorCount;
    if(wave_size) *wave_size = prop.warpSize;
    if(global_mem) *global_mem = prop.totalGlobalMem;
    if(name && name_len)
    {
        std::strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    return 0;
}

int
loom_kernel_count(void)
{
    return static_cast<int>(registry().size());
}

int
loom_kernel_info_get(int index, loom_kernel_info* out)
{
    if(int r = check_kernel(index)) return r;
    if(!out) return fail(LOOM_E_ARG, "null output");
    *out = registry()[index].info;
    return 0;
}

int
loom_kernel_dump(int kernel, char* buf, size_t len, size_t* needed)
{
    if(int r = check_kernel(kernel)) return r;
    const std::string text = registry()[kernel].dump();
    if(needed) *needed = text.size() + 1;
    if(buf && len)
    {
        const std::size_t n = std::min(len - 1, text.size());
        std::memcpy(buf, text.data(), n);
        buf[n] = 0;
    }
    return 0;
}

int
loom_kernel_resources(int kernel, int mode, int policy, int* vgprs, int* scratch_bytes)
{
    if(int r = ensure_device()) return r;
    if(int r = check_kernel(kernel)) return r;
    if(mode != LOOM_MODE_COMPACT && mode != LOOM_MODE_HYBRID)
        return fail(LOOM_E_ARG, "unknown mode %d", mode);
    const int all   = LOOM_RESOURCES_REGROUP | LOOM_RESOURCES_MEMORY | LOOM_RESOURCES_WORKGROUP |
                    LOOM_RESOURCES_WORKGROUP_8;
    const int flags = policy & all;
    policy &= ~all;
    if(policy < 0 || policy >= kNumPolicies) return fail(LOOM_E_ARG, "unknown policy %d", policy);
    if(int r = registry()[kernel].kernel_attrs(mode, policy, flags, vgprs, scratch_bytes))
        return fail(r, "hipFuncGetAttributes failed");
    return 0;
}

int
loom_kernel_register(int kernel, int r, char* name, size_t len, int* type, int* width, int* ntypes,
                     int* flags)
{
    if(int rc = check_kernel(kernel)) return rc;
    const char* nm = nullptr;
    std::size_t nl = 0;
    if(registry()[kernel].reg(r, &nm, &nl, type, width, ntypes, flags))
        return fail(LOOM_E_ARG, "kernel %s has no register %d", registry()[kernel].info.name, r);
    if(name && len)
    {
        const std::size_t n = std::min(len - 1, nl);
        std::memcpy(name, nm, n);
        name[n] = 0;
    }
    return 0;
}

int
loom_kernel_trace(int kernel, const void* const* inputs, uint64_t n, uint32_t max_steps,
                  uint64_t* rows, size_t rows_bytes, uint32_t* steps)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    if(int r = check_kernel(kernel)) return r;
    const KernelEntry& e = registry()[kernel];
    if(n >= (1ull << 20)) return fail(LOOM_E_CAPACITY, "at most 2^20-1 traced threads");
    if(max_steps == 0 || max_steps > (1u << 16))
        return fail(LOOM_E_ARG, "max_steps must be in [1, 65536]");
    if(n == 0) return 0;
    if(!inputs || !rows || !steps) return fail(LOOM_E_ARG, "null pointer argument");
    for(int k = 0; k < e.info.n_inputs; ++k)
        if(!inputs[k]) return fail(LOOM_E_ARG, "input column %d is null", k);
    // n < 2^20, max_steps <= 2^16, words <= 2 + max_regs: the product fits in 64 bits
    const std::uint64_t need =
        n * max_steps * std::uint64_t(LOOM_TRACE_WORDS(e.info.n_registers)) * sizeof(std::uint64_t);
    if(rows_bytes < need)
        return fail(LOOM_E_CAPACITY, "trace rows need %llu bytes, %zu given",
                    static_cast<unsigned long long>(need), rows_bytes);
    IoArgs io{};
    for(int k = 0; k < e.info.n_inputs; ++k) io.in[k] = inputs[k];
    LOOM_HIP(e.trace(io, n, max_steps, rows, steps, l.x->stream));
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_kernel_find(const char* name)
{
    if(!name) return -1;
    for(std::size_t i = 0; i < registry().size(); ++i)
        if(std::strcmp(registry()[i].info.name, name) == 0) return static_cast<int>(i);
    fail(LOOM_E_KERNEL, "no kernel named %s", name);
    return -1;
}

int
loom_malloc(void** ptr, size_t bytes)
{
    if(int r = ensure_device()) return r;
    if(!ptr) return fail(LOOM_E_ARG, "null pointer");
    LOOM_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int
loom_free(void* ptr)
{
    if(int r = ensure_device()) return r;
    LOOM_HIP(hipFree(ptr));
    return 0;
}

int
loom_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    LOOM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, l.x->stream));
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    LOOM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, l.x->stream));
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    LOOM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, l.x->stream));
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_memset(void* dst, int value, size_t bytes)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    LOOM_HIP(hipMemsetAsync(dst, value, bytes, l.x->stream));
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_ctx_fill(loom_context* ctx, void* dst, const void* pattern, size_t elem_bytes, uint64_t count)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    if(!count) return 0;
    if(!dst || !pattern) return fail(LOOM_E_ARG, "null pointer");
    hipStream_t s = l.x->stream;
    switch(elem_bytes)
    {
    case 1: LOOM_HIP(launch_fill<std::uint8_t>(dst, pattern, count, s)); break;
    case 2: LOOM_HIP(launch_fill<std::uint16_t>(dst, pattern, count, s)); break;
    case 4: LOOM_HIP(launch_fill<std::uint32_t>(dst, pattern, count, s)); break;
    case 8: LOOM_HIP(launch_fill<std::uint64_t>(dst, pattern, count, s)); break;
    default: return fail(LOOM_E_ARG, "element size %zu is not 1, 2, 4 or 8", elem_bytes);
    }
    LOOM_HIP(hipStreamSynchronize(s));
    return 0;
}

int
loom_fill(void* dst, const void* pattern, size_t elem_bytes, uint64_t count)
{
    return loom_ctx_fill(nullptr, dst, pattern, elem_bytes, count);
}

int
loom_host_alloc(void** ptr, size_t bytes)
{
    if(int r = ensure_device()) return r;
    if(!ptr) return fail(LOOM_E_ARG, "null pointer");
    LOOM_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return 0;
}

int
loom_host_free(void* ptr)
{
    if(int r = ensure_device()) return r;
    LOOM_HIP(hipHostFree(ptr));
    return 0;
}

int
loom_ctx_copy_async(loom_context* ctx, void* dst, const void* src, size_t bytes, int kind,
                    uint64_t* ticket)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    return copy_async_on(*l.x, dst, src, bytes, kind, ticket);
}

int
loom_copy_async(void* dst, const void* src, size_t bytes, int kind, uint64_t* ticket)
{
    return loom_ctx_copy_async(nullptr, dst, src, bytes, kind, ticket);
}

int
loom_ctx_copy_status(loom_context* ctx, uint64_t ticket, int* done)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    return copy_status_on(*l.x, ticket, done);
}

int
loom_copy_status(uint64_t ticket, int* done)
{
    return loom_ctx_copy_status(nullptr, ticket, done);
}

int
loom_ctx_copy_wait(loom_context* ctx, uint64_t ticket)
{
    hipEvent_t    ev{};
    loom_context* x = nullptr;
    {
        Locked l;
        if(int r = lock_context(ctx, l)) return r;
        x = l.x;
        if(ticket > x->last_ticket)
            return fail(LOOM_E_ARG, "ticket %llu was never issued",
                        static_cast<unsigned long long>(ticket));
        // outstanding tickets are consecutive: the one asked for, if still
        // outstanding, is at (ticket - oldest) in the queue
        if(!x->tickets.empty() && x->tickets.front().first <= ticket)
            ev = x->tickets[ticket - x->tickets.front().first].second;
    }
    // the event stays valid: it returns to the spare list only after it
    // completed, and a spare is re-recorded only by a later copy whose ticket
    // is larger, which this wait does not look at again
    if(ev) LOOM_HIP(hipEventSynchronize(ev));
    int done = 0;
    return loom_ctx_copy_status(ctx, ticket, &done);
}

int
loom_copy_wait(uint64_t ticket)
{
    return loom_ctx_copy_wait(nullptr, ticket);
}

int
loom_ctx_synchronize(loom_context* ctx)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    LOOM_HIP(hipStreamSynchronize(l.x->stream));
    return 0;
}

int
loom_synchronize(void)
{
    return loom_ctx_synchronize(nullptr);
}

int
loom_ctx_run(loom_context* ctx, int kernel, int mode, const void* const* inputs,
             void* const* outputs, uint64_t n, const loom_params* params, loom_run_stats* stats)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    return run_on(*l.x, kernel, mode, inputs, outputs, n, params, stats);
}

int
loom_run(int kernel, int mode, const void* const* inputs, void* const* outputs, uint64_t n,
         const loom_params* params, loom_run_stats* stats)
{
    return loom_ctx_run(nullptr, kernel, mode, inputs, outputs, n, params, stats);
}

int
loom_run_noalloc(int kernel, int mode, const void* const* inputs, void* const* outputs,
                 uint64_t n, const loom_params* params, loom_run_stats* stats)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    return run_on(*l.x, kernel, mode, inputs, outputs, n, params, stats, true);
}

int
loom_ctx_run_host(loom_context* ctx, int kernel, int mode, const void* const* inputs,
                  void* const* outputs, uint64_t n, const loom_params* params,
                  loom_run_stats* stats)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    return run_host_on(*l.x, kernel, mode, inputs, outputs, n, params, stats);
}

int
loom_run_host(int kernel, int mode, const void* const* inputs, void* const* outputs, uint64_t n,
              const loom_params* params, loom_run_stats* stats)
{
    return loom_ctx_run_host(nullptr, kernel, mode, inputs, outputs, n, params, stats);
}

int
loom_batch_set_count(void)
{
    return static_cast<int>(batch_registry().size());
}

int
loom_batch_set_find(const char* name)
{
    if(!name) return -1;
    for(std::size_t i = 0; i < batch_registry().size(); ++i)
        if(std::strcmp(batch_registry()[i].name, name) == 0) return static_cast<int>(i);
    fail(LOOM_E_KERNEL, "no batch set named %s", name);
    return -1;
}

int
loom_batch_set_members(int set, int* kernels, int max)
{
    if(set < 0 || set >= static_cast<int>(batch_registry().size()))
        return -fail(LOOM_E_KERNEL, "unknown batch set %d", set);
    const BatchEntry& b = batch_registry()[set];
    int               n = 0;
    for(std::uint32_t k = 0; k < b.size; ++k)
        for(std::size_t g = 0; g < registry().size(); ++g)
            if(registry()[g].info.hash == b.hash[k])
            {
                if(kernels && n < max) kernels[n] = static_cast<int>(g);
                ++n;
                break;
            }
    return n;
}

int
loom_ctx_batch_run(loom_context* ctx, int set, const loom_job* jobs, int njobs,
                   const loom_params* params, loom_run_stats* stats)
{
    Locked l;
    if(int r = lock_context(ctx, l)) return r;
    return batch_run_on(*l.x, set, jobs, njobs, params, stats);
}

int
loom_batch_run(int set, const loom_job* jobs, int njobs, const loom_params* params,
               loom_run_stats* stats)
{
    return loom_ctx_batch_run(nullptr, set, jobs, njobs, params, stats);
}

} // extern "C"

struct loom_graph
{
    hipGraph_t     graph{};
    hipGraphExec_t exec{};
    loom_context*  ctx  = nullptr; ///< the context whose stream and counters it uses
    JobDesc*       jobs = nullptr;
    BatchCtl*      ctl  = nullptr;
    loom_params    params{};
    bool           has_params = false;
    unsigned       grid       = 0;

    ~loom_graph()
    {
        if(exec) (void)hipGraphExecDestroy(exec);
        if(graph) (void)hipGraphDestroy(graph);
        if(jobs) (void)hipFree(jobs);
        if(ctl) (void)hipFree(ctl);
    }
};

extern "C" {

int
loom_batch_graph_create(int set, const loom_job* jobs, int njobs, const loom_params* params,
                        loom_graph** graph)
{
    Locked l;
    if(int r = lock_context(nullptr, l)) return r;
    loom_context& x = *l.x;
    if(!graph) return fail(LOOM_E_ARG, "null graph");
    if(int r = validate_params(params)) return r;
    if(params && params->register_file != LOOM_REGFILE_VGPR)
        return fail(LOOM_E_ARG, "batched launches keep registers in VGPRs");
    if(params && params->regroup == LOOM_REGROUP_WORKGROUP)
        return fail(LOOM_E_ARG, "batched launches do not regroup within workgroups");
    std::vector<JobDesc> jd;
    if(int r = batch_jobs(set, jobs, njobs, jd)) return r;
    // every early return below frees what was made so far (the destructor)
    auto g = std::make_unique<loom_graph>();
    g->ctx = &x;
    if(params)
    {
        g->params     = *params;
        g->has_params = true;
    }
    LOOM_HIP(hipMalloc(&g->jobs, sizeof(JobDesc) * kMaxJobs));
    LOOM_HIP(hipMalloc(&g->ctl, sizeof(BatchCtl)));
    LOOM_HIP(hipMemcpyAsync(g->jobs, jd.data(), sizeof(JobDesc) * jd.size(), hipMemcpyHostToDevice,
                            x.stream));
    LOOM_HIP(hipStreamSynchronize(x.stream));
    std::uint64_t threads = 0;
    for(const auto& j: jd) threads += j.end - j.begin;
    const BatchEntry& b = batch_registry()[set];
    const CtlArgs     c = make_ctl(params, threads, LOOM_MODE_HYBRID, x.err, x.stats);
    BatchArgs         a;
    a.jobs  = g->jobs;
    a.ctl   = g->ctl;
    a.njobs = static_cast<std::uint32_t>(njobs);
    g->grid = batch_grid(b, params, threads);
    // counters are cleared by a kernel node (no memset nodes in the graph)
    LOOM_HIP(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, x.stream, x.err, x.stats,
                       static_cast<std::uint32_t>(kStatWords));
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(kMaxJobs), 0, x.stream, a);
    const hipError_t le = b.launch[batch_policy(params)](c, a, g->grid, x.stream);
    const hipError_t ce = hipStreamEndCapture(x.stream, &g->graph);
    if(le != hipSuccess || ce != hipSuccess)
        return fail(LOOM_E_RUNTIME, "graph capture failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : ce));
    LOOM_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *graph = g.release();
    return 0;
}

int
loom_graph_launch(loom_graph* g, loom_run_stats* stats)
{
    if(!g) return fail(LOOM_E_ARG, "null graph");
    Locked l;
    if(int r = lock_context(g->ctx, l)) return r;
    loom_context& x = *l.x;
    if(stats) *stats = loom_run_stats{};
    LOOM_HIP(hipEventRecord(x.ev[0], x.stream));
    LOOM_HIP(hipGraphLaunch(g->exec, x.stream));
    LOOM_HIP(hipEventRecord(x.ev[1], x.stream));
    LOOM_HIP(hipEventSynchronize(x.ev[1]));
    float ms = 0;
    LOOM_HIP(hipEventElapsedTime(&ms, x.ev[0], x.ev[1]));
    if(stats)
    {
        stats->ms_total      = ms;
        stats->ms_main       = ms;
        stats->main_launches = 1;
        stats->passes        = 1;
        stats->grid_waves    = g->grid;
    }
    return finish_stats(x, stats, g->has_params ? &g->params : nullptr);
}

int
loom_graph_destroy(loom_graph* g)
{
    if(!g) return 0;
    {
        Locked l;
        if(lock_context(g->ctx, l) == 0) (void)hipStreamSynchronize(l.x->stream);
    }
    delete g;
    return 0;
}

} // extern "C"
