// -*- C++ -*-
/// @file trace.hpp
/// @brief Per-thread step trace: the state of each thread after every block
/// it runs, recorded on the GPU with the same lowered block code as the
/// compact and hybrid forms.
///
/// The reference's step-by-step inspector (src/vm/dbg.hpp:1011-1140) runs
/// one thread on the scalar backend and prints, after each block, the block
/// label and the values of the tracked variables (print_state,
/// dbg.hpp:920-935).  Here every lane runs its own thread one block at a
/// time and writes a row per block into a device buffer; the host formats
/// the rows.  A thread's trace does not depend on how a wave schedules its
/// lanes, so each lane simply runs the block at its own program counter (the
/// fold below predicates every block on `pc == B`).
///
/// Row of one step (2 + nr words of 64 bits):
///   word 0   block run (low 32 bits) | next block (high 32 bits, HALT = 0xFFFF)
///   word 1   error bits accumulated so far (LOOM_ERR_*)
///   word 2+r bits of register r after the block, zero-extended; registers
///            are zero before their first write.
#pragma once

#include "sched.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace loom::hipvm
{
/// @brief Words per trace row of kernel @p K.
template <typename K>
inline constexpr std::size_t kTraceWords = 2 + ir::table_of<K>.nr;

/// @brief Write the registers of @p r into @p row (words 2..).
template <typename K>
LOOM_DEVICE void
trace_registers(const Regs<K>& r, std::uint64_t* row)
{
    [&]<std::size_t... R>(std::index_sequence<R...>)
    { ((row[2 + R] = r.template raw<R>()), ...); }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief Trace threads [0, n): at most @p max_steps rows each, thread t's
/// rows at rows[t * max_steps * kTraceWords<K>]; steps[t] = rows written.
template <typename K>
__global__ __launch_bounds__(64) void
trace_kernel(IoArgs io, std::uint64_t n, std::uint32_t max_steps, std::uint64_t* rows,
             std::uint32_t* steps)
{
    sct_table_init<K>();
    const std::uint64_t tid = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(tid >= n) return;
    Regs<K> r;
    [&]<std::size_t... R>(std::index_sequence<R...>)
    { (r.template set_raw<R>(0), ...); }(std::make_index_sequence<ir::table_of<K>.nr>{});
    load_inputs<K>(r, io.in, tid);
    std::uint32_t  pc  = 0;
    std::uint32_t  err = 0;
    std::uint32_t  s   = 0;
    std::uint64_t* row = rows + tid * std::uint64_t(max_steps) * kTraceWords<K>;
    while(pc != HALT && s < max_steps)
    {
        const std::uint32_t b = pc;
        [&]<std::size_t... B>(std::index_sequence<B...>)
        {
            ((b == B ? exec_block<K, B>(r, pc, err) : void()), ...);
        }(std::make_index_sequence<ir::table_of<K>.nb>{});
        row[0] = b | (std::uint64_t(pc) << 32);
        row[1] = err;
        trace_registers<K>(r, row);
        row += kTraceWords<K>;
        ++s;
    }
    steps[tid] = s;
}
} // namespace loom::hipvm
