// -*- C++ -*-
/// @file stream.cpp
/// @brief The device-resident stream of the C ABI (loom_stream_* in
/// loom_amd.h): thread slots whose registers stay in HBM between steps,
/// exchange of finished threads for new inputs on the device.
///
/// Reference protocol (src/vm/simd.hpp:2797-2903, src/vm/bve.hpp:1208-1270,
/// driven by src/tests/stream_wrapper.hpp:33-80 and src/vm/mp.hpp:67-131):
/// set_inputs fills the slots, each step `s()` advances the live slots,
/// `sort()` reports whether an exchange is worth making, `exchange` swaps
/// the outputs of halted threads for new inputs in their slots, `done`
/// holds when no slot is live, and the outputs of the last threads are read
/// at the end.  Here the slots' registers never leave the device: a step is
/// one launch over the slots, an exchange copies only the exchanged rows'
/// inputs in and outputs out, and the slots that halted are found from a
/// ring the step kernel appends to.
///
/// Each stream owns its HIP stream (default priority), so streams driven by
/// different host threads run concurrently; calls on one stream are
/// serialised by its lock.
#include "runtime.hpp"

#include <algorithm>
#include <memory>

using namespace loom::hipvm;

struct loom_stream
{
    std::mutex         mu;
    int                kernel = -1;
    const KernelEntry* e      = nullptr;
    loom_params        params{};
    std::uint64_t      slots = 0;
    hipStream_t        stream{};
    StreamArgs         a{};
    char*              regs = nullptr;
    std::uint32_t*     pc   = nullptr;
    std::uint64_t*     tid  = nullptr;
    std::uint32_t*     ring = nullptr;
    StreamCtl*         ctl  = nullptr;
    StreamCtl*         host_ctl = nullptr; ///< pinned
    /// staging columns (device) of inputs in and outputs out, S rows each
    std::vector<char*> in_stage, out_stage;
    std::uint64_t*     ids_stage = nullptr;
    std::uint32_t*     err       = nullptr;
    std::uint64_t*     stats     = nullptr;
    std::uint32_t*     host_err  = nullptr; ///< pinned
    std::uint64_t*     host_stats = nullptr; ///< pinned
    hipEvent_t         ev[2]{};
    /// host view of the stream (exact after every call)
    std::uint64_t head = 0, tail = 0; ///< ring positions consumed / appended
    std::uint64_t live = 0;           ///< slots with a thread that has not halted
    std::uint64_t next_tid = 0;
    loom_stream_counters counters{};

    ~loom_stream()
    {
        if(stream) (void)hipStreamSynchronize(stream);
        for(void* p: {static_cast<void*>(regs), static_cast<void*>(pc), static_cast<void*>(tid),
                      static_cast<void*>(ring), static_cast<void*>(ctl),
                      static_cast<void*>(ids_stage), static_cast<void*>(err),
                      static_cast<void*>(stats)})
            if(p) (void)hipFree(p);
        for(char* p: in_stage)
            if(p) (void)hipFree(p);
        for(char* p: out_stage)
            if(p) (void)hipFree(p);
        if(host_ctl) (void)hipHostFree(host_ctl);
        if(host_err) (void)hipHostFree(host_err);
        if(host_stats) (void)hipHostFree(host_stats);
        for(auto ev: this->ev)
            if(ev) (void)hipEventDestroy(ev);
        if(stream) (void)hipStreamDestroy(stream);
    }
};

namespace
{
int
swap(loom_stream& s, std::uint32_t mode, std::uint64_t m)
{
    StreamIo x;
    for(std::size_t k = 0; k < s.in_stage.size(); ++k) x.in[k] = s.in_stage[k];
    for(std::size_t k = 0; k < s.out_stage.size(); ++k) x.out[k] = s.out_stage[k];
    x.ids      = s.ids_stage;
    x.head     = s.head;
    x.next_tid = s.next_tid;
    x.m        = static_cast<std::uint32_t>(m);
    x.mode     = mode;
    LOOM_HIP(s.e->stream_swap(s.a, x, s.stream));
    return 0;
}

/// Copy m rows of host input columns into the staging columns.
int
stage_inputs(loom_stream& s, const void* const* inputs, std::uint64_t m)
{
    for(int k = 0; k < s.e->info.n_inputs && m; ++k)
    {
        if(!inputs[k]) return fail(LOOM_E_ARG, "input column %d is null", k);
        const std::size_t b = m * type_size(s.e->info.input_type[k]);
        LOOM_HIP(hipMemcpyAsync(s.in_stage[k], inputs[k], b, hipMemcpyHostToDevice, s.stream));
        s.counters.bytes_h2d += b;
    }
    return 0;
}

/// Copy m rows of the staged outputs (and thread ids) to the host.
int
unstage_outputs(loom_stream& s, void* const* outputs, std::uint64_t* ids, std::uint64_t m)
{
    for(int k = 0; k < s.e->info.n_outputs && m; ++k)
    {
        if(!outputs[k]) return fail(LOOM_E_ARG, "output column %d is null", k);
        const std::size_t b = m * type_size(s.e->info.output_type[k]);
        LOOM_HIP(hipMemcpyAsync(outputs[k], s.out_stage[k], b, hipMemcpyDeviceToHost, s.stream));
        s.counters.bytes_d2h += b;
    }
    if(ids && m)
    {
        LOOM_HIP(hipMemcpyAsync(ids, s.ids_stage, m * sizeof(std::uint64_t), hipMemcpyDeviceToHost,
                                s.stream));
        s.counters.bytes_d2h += m * sizeof(std::uint64_t);
    }
    return 0;
}

int
check_stream(loom_stream* s)
{
    if(!s) return fail(LOOM_E_ARG, "null stream");
    return 0;
}
} // namespace

extern "C" {

int
loom_stream_create(int kernel, uint64_t slots, const loom_params* params, loom_stream** out)
{
    if(int r = ensure_device()) return r;
    if(int r = check_kernel(kernel)) return r;
    if(!out) return fail(LOOM_E_ARG, "null stream");
    if(slots == 0 || slots % 64 || slots > (1ull << 31))
        return fail(LOOM_E_ARG, "slots must be a positive multiple of 64, at most 2^31");
    if(int r = validate_params(params)) return r;
    if(params && params->register_file != LOOM_REGFILE_VGPR)
        return fail(LOOM_E_ARG, "a stream keeps its registers in device memory between steps and "
                                "in VGPRs during them");
    if(params && params->regroup == LOOM_REGROUP_WORKGROUP)
        return fail(LOOM_E_ARG, "a stream does not regroup within workgroups");
    auto s    = std::make_unique<loom_stream>();
    s->kernel = kernel;
    s->e      = &kernels()[kernel];
    if(params)
        s->params = *params;
    else
    {
        s->params.live_threshold = 0;
        s->params.reconv         = POLICY_SWEEP;
    }
    s->slots = slots;
    LOOM_HIP(hipSetDevice(device().id));
    LOOM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    const std::size_t state = std::max<std::uint32_t>(s->e->state_bytes, 4);
    LOOM_HIP(hipMalloc(&s->regs, slots * state));
    LOOM_HIP(hipMalloc(&s->pc, slots * sizeof(std::uint32_t)));
    LOOM_HIP(hipMalloc(&s->tid, slots * sizeof(std::uint64_t)));
    LOOM_HIP(hipMalloc(&s->ring, slots * sizeof(std::uint32_t)));
    LOOM_HIP(hipMalloc(&s->ctl, sizeof(StreamCtl)));
    LOOM_HIP(hipMalloc(&s->ids_stage, slots * sizeof(std::uint64_t)));
    LOOM_HIP(hipMalloc(&s->err, sizeof(std::uint32_t)));
    LOOM_HIP(hipMalloc(&s->stats, sizeof(std::uint64_t) * kStatWords));
    LOOM_HIP(hipHostMalloc(&s->host_ctl, sizeof(StreamCtl), hipHostMallocDefault));
    LOOM_HIP(hipHostMalloc(&s->host_err, sizeof(std::uint32_t), hipHostMallocDefault));
    LOOM_HIP(hipHostMalloc(&s->host_stats, sizeof(std::uint64_t) * kStatWords, hipHostMallocDefault));
    for(auto& ev: s->ev) LOOM_HIP(hipEventCreate(&ev));
    for(int k = 0; k < s->e->info.n_inputs; ++k)
    {
        s->in_stage.push_back(nullptr);
        LOOM_HIP(hipMalloc(&s->in_stage.back(), slots * type_size(s->e->info.input_type[k])));
    }
    for(int k = 0; k < s->e->info.n_outputs; ++k)
    {
        s->out_stage.push_back(nullptr);
        LOOM_HIP(hipMalloc(&s->out_stage.back(), slots * type_size(s->e->info.output_type[k])));
    }
    s->a.regs  = s->regs;
    s->a.pc    = s->pc;
    s->a.tid   = s->tid;
    s->a.ring  = s->ring;
    s->a.ctl   = s->ctl;
    s->a.slots = slots;
    // every slot empty, the ring empty
    LOOM_HIP(hipMemsetAsync(s->ctl, 0, sizeof(StreamCtl), s->stream));
    if(int r = swap(*s, SWAP_INIT, 0)) return r;
    LOOM_HIP(hipStreamSynchronize(s->stream));
    *out = s.release();
    return 0;
}

int
loom_stream_destroy(loom_stream* s)
{
    if(!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
    }
    delete s;
    return 0;
}

int
loom_stream_set_inputs(loom_stream* s, const void* const* inputs, uint64_t n, uint64_t* taken)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(n && !inputs) return fail(LOOM_E_ARG, "null column array");
    const std::uint64_t m = std::min<std::uint64_t>(n, s->slots);
    if(int r = stage_inputs(*s, inputs, m)) return r;
    s->head = s->tail = 0;
    s->next_tid       = 0;
    LOOM_HIP(hipMemsetAsync(s->ctl, 0, sizeof(StreamCtl), s->stream));
    if(int r = swap(*s, SWAP_INIT, m)) return r;
    LOOM_HIP(hipStreamSynchronize(s->stream));
    s->live     = m;
    s->next_tid = m;
    s->counters.rows_in += m;
    if(taken) *taken = m;
    return 0;
}

int
loom_stream_step(loom_stream* s, loom_run_stats* stats)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(stats) *stats = loom_run_stats{};
    const loom_params& p = s->params;
    CtlArgs            c = make_ctl(&p, s->live, LOOM_MODE_HYBRID, s->err, s->stats);
    LOOM_HIP(hipMemsetAsync(s->err, 0, sizeof(std::uint32_t), s->stream));
    if(c.stats) LOOM_HIP(hipMemsetAsync(s->stats, 0, sizeof(std::uint64_t) * kStatWords, s->stream));
    LOOM_HIP(hipEventRecord(s->ev[0], s->stream));
    if(s->live) LOOM_HIP(s->e->stream_step[p.reconv](s->a, c, s->stream));
    LOOM_HIP(hipEventRecord(s->ev[1], s->stream));
    LOOM_HIP(hipMemcpyAsync(s->host_ctl, s->ctl, sizeof(StreamCtl), hipMemcpyDeviceToHost, s->stream));
    LOOM_HIP(hipMemcpyAsync(s->host_err, s->err, sizeof(std::uint32_t), hipMemcpyDeviceToHost,
                            s->stream));
    if(c.stats)
        LOOM_HIP(hipMemcpyAsync(s->host_stats, s->stats, sizeof(std::uint64_t) * kStatWords,
                                hipMemcpyDeviceToHost, s->stream));
    LOOM_HIP(hipStreamSynchronize(s->stream));
    const std::uint64_t halted = s->host_ctl->tail.v - s->tail;
    s->tail                    = s->host_ctl->tail.v;
    s->live -= std::min(halted, s->live);
    ++s->counters.steps;
    float ms = 0;
    LOOM_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    if(stats)
    {
        stats->ms_total      = ms;
        stats->ms_main       = ms;
        stats->main_launches = 1;
        stats->passes        = 1;
        stats->grid_waves    = static_cast<std::uint32_t>(s->slots / 64);
        stats->errors        = *s->host_err;
        if(c.stats) sum_counters(s->host_stats, stats->counters);
    }
    if(*s->host_err) return fail(LOOM_E_RUNTIME, "kernel raised error flags 0x%x", *s->host_err);
    return 0;
}

int
loom_stream_sort(loom_stream* s, int* exchange_available)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(!exchange_available) return fail(LOOM_E_ARG, "null result");
    const std::uint64_t halted = s->tail - s->head;
    // a wave's worth of finished slots, or every finished slot once none is live
    *exchange_available = halted >= 64 || (halted > 0 && s->live == 0);
    return 0;
}

int
loom_stream_exchange(loom_stream* s, const void* const* inputs, uint64_t n, void* const* outputs,
                     uint64_t* ids, uint64_t* exchanged)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(exchanged) *exchanged = 0;
    const std::uint64_t m = std::min<std::uint64_t>(n, s->tail - s->head);
    if(m == 0) return 0;
    if(!inputs || !outputs) return fail(LOOM_E_ARG, "null column array");
    if(int r = stage_inputs(*s, inputs, m)) return r;
    if(int r = swap(*s, SWAP_EXCHANGE, m)) return r;
    if(int r = unstage_outputs(*s, outputs, ids, m)) return r;
    LOOM_HIP(hipStreamSynchronize(s->stream));
    s->head += m;
    s->live += m;
    s->next_tid += m;
    s->counters.exchanges += 1;
    s->counters.rows_in += m;
    s->counters.rows_out += m;
    if(exchanged) *exchanged = m;
    return 0;
}

int
loom_stream_drain(loom_stream* s, void* const* outputs, uint64_t* ids, uint64_t cap,
                  uint64_t* drained)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(drained) *drained = 0;
    const std::uint64_t m = std::min<std::uint64_t>(cap, s->tail - s->head);
    if(m == 0) return 0;
    if(!outputs) return fail(LOOM_E_ARG, "null column array");
    if(int r = swap(*s, SWAP_DRAIN, m)) return r;
    if(int r = unstage_outputs(*s, outputs, ids, m)) return r;
    LOOM_HIP(hipStreamSynchronize(s->stream));
    s->head += m;
    s->counters.rows_out += m;
    if(drained) *drained = m;
    return 0;
}

int
loom_stream_state(loom_stream* s, uint64_t* live, uint64_t* halted)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(live) *live = s->live;
    if(halted) *halted = s->tail - s->head;
    return 0;
}

int
loom_stream_counters_get(loom_stream* s, loom_stream_counters* out)
{
    if(int r = check_stream(s)) return r;
    std::lock_guard<std::mutex> lk(s->mu);
    if(!out) return fail(LOOM_E_ARG, "null result");
    *out = s->counters;
    return 0;
}

} // extern "C"
