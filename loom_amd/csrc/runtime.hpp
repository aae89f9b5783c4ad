// -*- C++ -*-
/// @file runtime.hpp
/// @brief Internal interface of the host runtime shared by its translation
/// units (runtime.cpp, stream.cpp, serve.cpp): the kernel registry, the
/// device, error reporting, and the execution context.
#pragma once

#include "registry.hpp"

#include <loom_amd.h>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <mutex>
#include <vector>

namespace loom::hipvm
{
/// @brief The registry of lowered kernels and batch sets.  Entries are added
/// while the library (and modules linked with it) are loaded and are read
/// only afterwards, so lookups need no lock.
const std::vector<KernelEntry>& kernels();
const std::vector<BatchEntry>&  batch_sets();

/// @brief The device the library was initialised on (read only after
/// loom_init).
struct Device
{
    bool ready = false;
    int  id    = 0;
    int  cus   = 0;
};
const Device& device();

/// @brief Record a failure for loom_last_error (per host thread) and return
/// @p code.
int fail(int code, const char* fmt, ...);

#define LOOM_HIP(expr)                                                                 \
    do                                                                                 \
    {                                                                                  \
        const hipError_t e_ = (expr);                                                  \
        if(e_ != hipSuccess)                                                           \
            return ::loom::hipvm::fail(static_cast<int>(e_), "%s: %s", #expr,          \
                                       hipGetErrorString(e_));                         \
    } while(0)

int ensure_device();
int check_kernel(int k);
int type_size(int t);
int validate_params(const loom_params* params);

/// @brief Run controls of a launch from @p p (NULL: the defaults of
/// loom_params), with the error flags and counters of @p err / @p stats.
CtlArgs make_ctl(const loom_params* p, std::uint64_t n, int mode, std::uint32_t* err,
                 std::uint64_t* stats);

/// @brief Sum the counter copies of a run into @p out.
void sum_counters(const std::uint64_t* host_stats, std::uint64_t* out);
} // namespace loom::hipvm

/// @brief An execution context: one HIP stream for kernels and one for
/// background copies, the run's error flags and counters, and the
/// workspaces of hybrid and batched runs.  Calls on one context are
/// serialised by its lock; calls on different contexts run concurrently
/// (the GPU overlaps the kernels of their streams).
struct loom_context
{
    std::mutex     mu;
    hipStream_t    stream{};
    hipStream_t    copy{};
    std::uint32_t* err        = nullptr; ///< device error flags
    std::uint64_t* stats      = nullptr; ///< device counters (kStatWords)
    std::uint32_t* host_err   = nullptr; ///< pinned
    std::uint64_t* host_stats = nullptr; ///< pinned
    loom::hipvm::HybridCtl* host_ctl = nullptr; ///< pinned mirror of a control block
    hipEvent_t     ev[4]{};
    std::vector<hipEvent_t> pass_ev;
    loom::hipvm::JobDesc*  jobs = nullptr; ///< device job table of direct batch runs
    loom::hipvm::BatchCtl* bctl = nullptr;
    /// background copies: tickets of outstanding copies, oldest first
    std::deque<std::pair<std::uint64_t, hipEvent_t>> tickets;
    std::vector<hipEvent_t> spare_events;
    std::uint64_t           last_ticket = 0;

    struct Workspace
    {
        loom::hipvm::HybridCtl* ctl   = nullptr;
        char*                   q[2]  = {nullptr, nullptr};
        std::size_t             bytes = 0; ///< per side
    };
    std::vector<Workspace> ws; ///< per kernel index

    /// input cursor of workgroup-regrouping runs
    std::uint64_t* wgr_next = nullptr;

    /// register file of compact runs in device memory, grown on demand
    char*       memregs       = nullptr;
    std::size_t memregs_bytes = 0;

    /// device columns of loom_run_host, grown on demand
    std::vector<void*>       hostcol_ptr;
    std::vector<std::size_t> hostcol_bytes;
};

namespace loom::hipvm
{
/// @brief The context of the legacy entry points (created by loom_init).
loom_context* default_context();
} // namespace loom::hipvm
