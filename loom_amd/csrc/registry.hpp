// -*- C++ -*-
/// @file registry.hpp
/// @brief Registration of lowered kernels with the runtime.
///
/// A translation unit compiled for gfx950 lowers a kernel and instantiates its
/// device kernels with `LOOM_HIP_REGISTER("name", K)`.  The entry records the
/// kernel table summary and launcher functions; the C ABI finds kernels by
/// index or name.
#pragma once

#include "device/batch.hpp"
#include "device/stream.hpp"
#include "device/trace.hpp"
#include "device/wgr.hpp"

#include <loom_amd.h>

#include <hip/hip_runtime.h>

#include <vm/dump.hpp>

#include <cstdint>
#include <cstring>
#include <string>

namespace loom::hipvm
{
struct KernelEntry
{
    loom_kernel_info info{};
    /// launchers per reconvergence policy (POLICY_SUCCESSOR_MAJOR, POLICY_TOPO,
    /// POLICY_SWEEP)
    hipError_t (*compact[kNumPolicies])(const IoArgs&, const CtlArgs&, hipStream_t) = {};
    /// the compact kernel with its register file in device memory (sweep
    /// policy; MemArgs.rows >= the run's threads)
    hipError_t (*compact_mem)(const IoArgs&, const CtlArgs&, const MemArgs&, hipStream_t) = nullptr;
    /// the workgroup-regrouping kernel (successor-major, sweep; workgroups
    /// of `waves` = 16 or 8 waves) and its resident workgroups per CU
    hipError_t (*wgr[2])(const IoArgs&, const CtlArgs&, const WgrArgs&, int waves,
                         unsigned grid, hipStream_t) = {};
    int (*wgr_occupancy[2])(int waves) = {};
    hipError_t (*hybrid_pass[kNumPolicies])(const IoArgs&, const CtlArgs&, const HybridArgs&,
                                 unsigned grid, hipStream_t)           = {};
    /// resident 64-thread workgroups per CU of the hybrid pass kernel
    /// (with or without the regroup paths)
    int (*hybrid_occupancy[kNumPolicies])(bool regroup)                = {};
    /// registers and scratch of the device kernels (0: compact, 1: hybrid)
    int (*kernel_attrs)(int mode, int policy, int flags, int* vgprs, int* scratch) = nullptr;
    std::string (*dump)() = nullptr;
    /// step trace of threads [0, n) (trace_kernel)
    hipError_t (*trace)(const IoArgs&, std::uint64_t n, std::uint32_t max_steps,
                        std::uint64_t* rows, std::uint32_t* steps, hipStream_t) = nullptr;
    /// device-resident stream: one step over every live slot (per policy),
    /// and the initialise / exchange / drain swap of slot rows
    hipError_t (*stream_step[kNumPolicies])(const StreamArgs&, const CtlArgs&, hipStream_t) = {};
    hipError_t (*stream_swap)(const StreamArgs&, const StreamIo&, hipStream_t) = nullptr;
    /// register bytes of one stream slot (sum of the register widths)
    std::uint32_t state_bytes = 0;
    /// name, first type, width, type count and LOOM_REG_* flags of register r
    /// (returns 0 when r < nr)
    int (*reg)(int r, const char** name, std::size_t* name_len, int* type, int* width,
               int* ntypes, int* flags) = nullptr;
};

/// @brief Add @p e to the process-wide registry; returns true.
bool register_kernel(const KernelEntry& e);

/// @brief A compiled batch set: kernels that one persistent launch can mix.
struct BatchEntry
{
    char          name[64]{};
    std::uint32_t size = 0;
    std::uint64_t hash[kMaxJobs]{}; ///< structural hashes of the members
    /// launchers and occupancy per reconvergence policy
    hipError_t (*launch[kNumPolicies])(const CtlArgs&, const BatchArgs&, unsigned grid,
                                       hipStream_t) = {};
    int (*occupancy[kNumPolicies])() = {};
};

bool register_batch(const BatchEntry& e);

template <int P, typename... Ks>
hipError_t
launch_batch(const CtlArgs& c, const BatchArgs& a, unsigned grid, hipStream_t s)
{
    hipLaunchKernelGGL((batch_kernel<P, Ks...>), dim3(grid), dim3(64), 0, s, c, a);
    return hipGetLastError();
}

template <int P, typename... Ks>
int
batch_occupancy()
{
    int nblk = 0;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, batch_kernel<P, Ks...>, 64, 0) !=
       hipSuccess)
        return 1;
    return nblk > 0 ? nblk : 1;
}

template <typename... Ks>
BatchEntry
make_batch(const char* name)
{
    BatchEntry e;
    std::strncpy(e.name, name, sizeof(e.name) - 1);
    e.size        = sizeof...(Ks);
    std::size_t i = 0;
    ((e.hash[i++] = ir::table_of<Ks>.hash), ...);
    e.launch[0]    = &launch_batch<POLICY_SUCCESSOR_MAJOR, Ks...>;
    e.launch[1]    = &launch_batch<POLICY_TOPO, Ks...>;
    e.launch[2]    = &launch_batch<POLICY_SWEEP, Ks...>;
    e.occupancy[0] = &batch_occupancy<POLICY_SUCCESSOR_MAJOR, Ks...>;
    e.occupancy[1] = &batch_occupancy<POLICY_TOPO, Ks...>;
    e.occupancy[2] = &batch_occupancy<POLICY_SWEEP, Ks...>;
    return e;
}

template <typename K, int P>
hipError_t
launch_compact(const IoArgs& io, const CtlArgs& c, hipStream_t s)
{
    if(c.n == 0) return hipSuccess;
    const std::uint64_t blocks = (c.n + 255) / 256;
    // a run without statistics takes the kernel without per-lane work
    // counts where the kernel has small blocks (the same kernel otherwise)
    if constexpr(has_small_blocks<K>)
    {
        if(!c.stats)
        {
            hipLaunchKernelGGL((compact_kernel<K, P, false>), dim3(static_cast<unsigned>(blocks)),
                               dim3(256), 0, s, io, c);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((compact_kernel<K, P>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       io, c);
    return hipGetLastError();
}

template <typename K>
hipError_t
launch_compact_mem(const IoArgs& io, const CtlArgs& c, const MemArgs& m, hipStream_t s)
{
    if(c.n == 0) return hipSuccess;
    const std::uint64_t blocks = (c.n + 255) / 256;
    hipLaunchKernelGGL((compact_mem_kernel<K, POLICY_SWEEP>), dim3(static_cast<unsigned>(blocks)),
                       dim3(256), 0, s, io, c, m);
    return hipGetLastError();
}

template <typename K, int P>
hipError_t
launch_wgr(const IoArgs& io, const CtlArgs& c, const WgrArgs& g, int waves, unsigned grid,
           hipStream_t s)
{
    if(c.n == 0) return hipSuccess;
    if constexpr(has_small_blocks<K>)
    {
        if(!c.stats)
        {
            if(waves == 8)
                hipLaunchKernelGGL((wgr_kernel<K, P, 8, false>), dim3(grid), dim3(64 * 8), 0, s, io,
                                   c, g);
            else
                hipLaunchKernelGGL((wgr_kernel<K, P, 16, false>), dim3(grid), dim3(64 * 16), 0, s,
                                   io, c, g);
            return hipGetLastError();
        }
    }
    if(waves == 8)
        hipLaunchKernelGGL((wgr_kernel<K, P, 8>), dim3(grid), dim3(64 * 8), 0, s, io, c, g);
    else
        hipLaunchKernelGGL((wgr_kernel<K, P, 16>), dim3(grid), dim3(64 * 16), 0, s, io, c, g);
    return hipGetLastError();
}

template <typename K, int P>
int
wgr_occupancy(int waves)
{
    int              nblk = 0;
    const hipError_t e =
        waves == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, wgr_kernel<K, P, 8>,
                                                                  64 * 8, 0) :
                     hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, wgr_kernel<K, P, 16>,
                                                                  64 * 16, 0);
    if(e != hipSuccess) return 1;
    return nblk > 0 ? nblk : 1;
}

/// Runs that regroup launch the pass kernel with the queue paths compiled
/// in; the others the kernel without them, and those of them that collect
/// no statistics, for a kernel with small blocks, the one without per-lane
/// work counts.
template <typename K, int P>
hipError_t
launch_hybrid_pass(const IoArgs& io, const CtlArgs& c, const HybridArgs& h, unsigned grid,
                   hipStream_t s)
{
    if(c.regroup)
        hipLaunchKernelGGL((hybrid_pass_kernel<K, P, true>), dim3(grid), dim3(64), 0, s, io, c, h);
    else if(has_small_blocks<K> && !c.stats)
        hipLaunchKernelGGL((hybrid_pass_kernel<K, P, false, !has_small_blocks<K>>), dim3(grid),
                           dim3(64), 0, s, io, c, h);
    else
        hipLaunchKernelGGL((hybrid_pass_kernel<K, P, false>), dim3(grid), dim3(64), 0, s, io, c, h);
    return hipGetLastError();
}

template <typename K, int P>
int
hybrid_occupancy(bool regroup)
{
    int nblk = 0;
    const hipError_t e =
        regroup ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, hybrid_pass_kernel<K, P, true>,
                                                               64, 0) :
                  hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, hybrid_pass_kernel<K, P, false>,
                                                               64, 0);
    if(e != hipSuccess) return 1;
    return nblk > 0 ? nblk : 1;
}

/// @brief Registers and scratch of the compiled kernel of @p mode and
/// @p policy (LOOM_MODE_COMPACT: the predicated kernel, with its register
/// file in memory under LOOM_RESOURCES_MEMORY; LOOM_MODE_HYBRID: the pass
/// kernel, with the regroup paths under LOOM_RESOURCES_REGROUP).
template <typename K>
int
kernel_attrs(int mode, int policy, int flags, int* vgprs, int* scratch)
{
    const bool  regroup = (flags & LOOM_RESOURCES_REGROUP) != 0;
    const bool  memory  = (flags & LOOM_RESOURCES_MEMORY) != 0;
    const bool  wgr     = (flags & LOOM_RESOURCES_WORKGROUP) != 0;
    const bool  wgr8    = (flags & LOOM_RESOURCES_WORKGROUP_8) != 0;
    const void* f = nullptr;
    auto        pick = [&]<int P>()
    {
        constexpr int WP = P == POLICY_SUCCESSOR_MAJOR ? POLICY_SUCCESSOR_MAJOR : POLICY_SWEEP;
        if(mode == LOOM_MODE_HYBRID && wgr8)
            f = reinterpret_cast<const void*>(wgr_kernel<K, WP, 8>);
        else if(mode == LOOM_MODE_HYBRID && wgr)
            f = reinterpret_cast<const void*>(wgr_kernel<K, WP, 16>);
        else if(mode == LOOM_MODE_COMPACT && memory)
            f = reinterpret_cast<const void*>(compact_mem_kernel<K, POLICY_SWEEP>);
        else if(mode == LOOM_MODE_COMPACT)
            f = reinterpret_cast<const void*>(compact_kernel<K, P>);
        else if(regroup)
            f = reinterpret_cast<const void*>(hybrid_pass_kernel<K, P, true>);
        else
            f = reinterpret_cast<const void*>(hybrid_pass_kernel<K, P, false>);
    };
    if(policy == POLICY_SUCCESSOR_MAJOR) pick.template operator()<POLICY_SUCCESSOR_MAJOR>();
    else if(policy == POLICY_TOPO) pick.template operator()<POLICY_TOPO>();
    else pick.template operator()<POLICY_SWEEP>();
    hipFuncAttributes a{};
    const hipError_t  e = hipFuncGetAttributes(&a, f);
    if(e != hipSuccess) return static_cast<int>(e);
    if(vgprs) *vgprs = a.numRegs;
    if(scratch) *scratch = static_cast<int>(a.localSizeBytes);
    return 0;
}

template <typename K, int P>
hipError_t
launch_stream_step(const StreamArgs& a, const CtlArgs& c, hipStream_t s)
{
    if(a.slots == 0) return hipSuccess;
    hipLaunchKernelGGL((stream_step_kernel<K, P>), dim3(static_cast<unsigned>(a.slots / 64)),
                       dim3(64), 0, s, a, c);
    return hipGetLastError();
}

template <typename K>
hipError_t
launch_stream_swap(const StreamArgs& a, const StreamIo& x, hipStream_t s)
{
    const std::uint64_t rows = x.mode == SWAP_INIT ? a.slots : x.m;
    if(rows == 0) return hipSuccess;
    hipLaunchKernelGGL((stream_swap_kernel<K>), dim3(static_cast<unsigned>((rows + 255) / 256)),
                       dim3(256), 0, s, a, x);
    return hipGetLastError();
}

template <typename K>
hipError_t
launch_trace(const IoArgs& io, std::uint64_t n, std::uint32_t max_steps, std::uint64_t* rows,
             std::uint32_t* steps, hipStream_t s)
{
    if(n == 0) return hipSuccess;
    const std::uint64_t blocks = (n + 63) / 64;
    hipLaunchKernelGGL((trace_kernel<K>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, io,
                       n, max_steps, rows, steps);
    return hipGetLastError();
}

template <typename K>
int
reg_of(int r, const char** name, std::size_t* name_len, int* type, int* width, int* ntypes,
       int* flags)
{
    constexpr const auto& T = ir::table_of<K>;
    if(r < 0 || static_cast<std::size_t>(r) >= T.nr) return 1;
    const ir::Reg& g = T.reg[r];
    if(name) *name = g.name.data();
    if(name_len) *name_len = g.name.size();
    if(type) *type = static_cast<int>(g.ty);
    if(width) *width = g.width;
    if(ntypes) *ntypes = g.ntypes;
    if(flags)
        *flags = (g.input ? LOOM_REG_INPUT : 0) | (g.output ? LOOM_REG_OUTPUT : 0) |
                 (g.written ? LOOM_REG_WRITTEN : 0);
    return 0;
}

/// @brief Text listing of the lowered table of @p K.
template <typename K>
std::string
dump_of()
{
    return ir::dump_table(ir::table_of<K>);
}

template <typename K>
KernelEntry
make_entry(const char* name)
{
    constexpr const auto& T = ir::table_of<K>;
    static_assert(T.nin <= kMaxIO && T.nout <= kMaxIO, "too many kernel inputs/outputs");
    KernelEntry e;
    std::strncpy(e.info.name, name, sizeof(e.info.name) - 1);
    e.info.hash        = T.hash;
    e.info.n_inputs    = static_cast<std::int32_t>(T.nin);
    e.info.n_outputs   = static_cast<std::int32_t>(T.nout);
    for(std::size_t k = 0; k < T.nin; ++k)
        e.info.input_type[k] = static_cast<std::int32_t>(T.in_ty[k]);
    for(std::size_t k = 0; k < T.nout; ++k)
        e.info.output_type[k] = static_cast<std::int32_t>(T.out_ty[k]);
    e.info.n_blocks        = static_cast<std::int32_t>(T.nb);
    e.info.n_registers     = static_cast<std::int32_t>(T.nr);
    e.info.index_bytes     = T.index_bytes;
    e.info.state_size      = static_cast<std::uint32_t>(T.state_size);
    e.info.queue_row_bytes = QueueLayout<K>::total;
    for(std::size_t b = 0; b < T.nb; ++b)
    {
        e.info.block_ops[b]        = T.block[b].ops;
        e.info.block_live_bytes[b] = QueueLayout<K>::live_bytes(b);
        e.info.block_flops[b]      = T.block[b].flops;
    }
    e.compact[0]          = &launch_compact<K, POLICY_SUCCESSOR_MAJOR>;
    e.compact[1]          = &launch_compact<K, POLICY_TOPO>;
    e.compact[2]          = &launch_compact<K, POLICY_SWEEP>;
    e.compact_mem         = &launch_compact_mem<K>;
    e.wgr[0]              = &launch_wgr<K, POLICY_SUCCESSOR_MAJOR>;
    e.wgr[1]              = &launch_wgr<K, POLICY_SWEEP>;
    e.wgr_occupancy[0]    = &wgr_occupancy<K, POLICY_SUCCESSOR_MAJOR>;
    e.wgr_occupancy[1]    = &wgr_occupancy<K, POLICY_SWEEP>;
    e.hybrid_pass[0]      = &launch_hybrid_pass<K, POLICY_SUCCESSOR_MAJOR>;
    e.hybrid_pass[1]      = &launch_hybrid_pass<K, POLICY_TOPO>;
    e.hybrid_pass[2]      = &launch_hybrid_pass<K, POLICY_SWEEP>;
    e.hybrid_occupancy[0] = &hybrid_occupancy<K, POLICY_SUCCESSOR_MAJOR>;
    e.hybrid_occupancy[1] = &hybrid_occupancy<K, POLICY_TOPO>;
    e.hybrid_occupancy[2] = &hybrid_occupancy<K, POLICY_SWEEP>;
    e.stream_step[0]      = &launch_stream_step<K, POLICY_SUCCESSOR_MAJOR>;
    e.stream_step[1]      = &launch_stream_step<K, POLICY_TOPO>;
    e.stream_step[2]      = &launch_stream_step<K, POLICY_SWEEP>;
    e.stream_swap         = &launch_stream_swap<K>;
    e.state_bytes         = StateLayout<K>::bytes;
    e.kernel_attrs     = &kernel_attrs<K>;
    e.dump             = &dump_of<K>;
    e.trace            = &launch_trace<K>;
    e.reg              = &reg_of<K>;
    return e;
}
} // namespace loom::hipvm

#define LOOM_REG_CAT(a, b) LOOM_REG_CAT_I(a, b)
#define LOOM_REG_CAT_I(a, b) a##b
#define LOOM_HIP_REGISTER_I(LINE, NAME, ...)                              \
    static const bool LOOM_REG_CAT(loom_registered_, LINE) =               \
        ::loom::hipvm::register_kernel(::loom::hipvm::make_entry<__VA_ARGS__>(NAME))

/// @brief Compile the hybrid pass kernel without the regroup paths of kernel
/// type `...` for at least @p W waves per SIMD (before its LOOM_HIP_REGISTER,
/// at namespace scope).
#define LOOM_HIP_WAVES_HINT(W, ...)                                                   \
    template <>                                                                        \
    inline constexpr int ::loom::hipvm::waves_hint<__VA_ARGS__> = (W)

/// @brief The same for the pass kernel with the regroup paths.
#define LOOM_HIP_REGROUP_WAVES_HINT(W, ...)                                           \
    template <>                                                                        \
    inline constexpr int ::loom::hipvm::regroup_waves_hint<__VA_ARGS__> = (W)

/// @brief Lower kernel type `...` and register it under @p NAME.
#define LOOM_HIP_REGISTER(NAME, ...) LOOM_HIP_REGISTER_I(__COUNTER__, NAME, __VA_ARGS__)

/// @brief Compile one persistent kernel mixing the kernel types `...` and
/// register it as batch set @p NAME.
#define LOOM_HIP_REGISTER_BATCH(NAME, ...)                                   \
    static const bool LOOM_REG_CAT(loom_batch_registered_, __COUNTER__) =    \
        ::loom::hipvm::register_batch(::loom::hipvm::make_batch<__VA_ARGS__>(NAME))
