// -*- C++ -*-
/// @file memregs.hpp
/// @brief A register file in device memory: one column per register, one
/// row per thread (structure of arrays, the reference's bve::Memory layout,
/// src/vm/bve.hpp:494-693).  A block run on such a lane loads the registers
/// the block reads before writing them, runs on VGPR copies, and stores the
/// registers it writes that are live after it; VGPRs hold one block's
/// working set instead of the whole register file, so a kernel's size is
/// bounded by memory rather than by the 512 VGPRs of a lane.
#pragma once

#include "exec.hpp"

namespace loom::hipvm
{
/// @brief Byte offsets of the register columns per row (before * rows).  A
/// column holds the register's cell: 8 bytes for a 64-bit register, 4 for
/// every narrower one (a 1- or 2-byte register lives in a 32-bit cell, as in
/// the VGPR register file, and is read and written as one).
template <typename K>
struct CellLayout
{
    static constexpr auto off = []
    {
        std::array<std::uint32_t, ir::max_regs + 1> o{};
        for(std::size_t r = 0; r < ir::table_of<K>.nr; ++r)
            o[r + 1] = o[r] + (ir::table_of<K>.reg[r].width == 8 ? 8u : 4u);
        return o;
    }();
    static constexpr std::uint32_t bytes = off[ir::table_of<K>.nr];
};

/// @brief Registers live after block @p B: those live into its successors,
/// and the outputs (a multiway jump whose index is beyond its table halts).
template <typename K, std::size_t B>
inline constexpr auto live_out = []
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];
    auto                o   = ir::table_of<K>.outputs;
    for(std::size_t j = 0; j < blk.ns; ++j)
    {
        const auto s = ir::table_of<K>.succ[blk.s0 + j];
        if(s != ir::halt16) o = o | ir::table_of<K>.block[s].live_in;
    }
    return o;
}();

/// @brief This lane's row of a register file in device memory.  A column's
/// address is wave-uniform (scalar registers) and the lane's offset in it is
/// its row times the cell size, so an access is one load or store with a
/// scalar base and a 32-bit lane offset.
template <typename K>
struct MemRegs
{
    char*         base = nullptr; ///< column 0 (wave-uniform)
    std::uint64_t rows = 0;       ///< elements per column (wave-uniform)
    std::uint32_t row  = 0;       ///< this lane's row

    template <std::size_t R>
    LOOM_DEVICE std::uint64_t
    load() const
    {
        const char* col = base + rows * LOOM_CX(CellLayout<K>::off[R]);
        if constexpr(ir::table_of<K>.reg[R].width == 8)
            return reinterpret_cast<const std::uint64_t*>(col)[row];
        else
            return reinterpret_cast<const std::uint32_t*>(col)[row];
    }

    template <std::size_t R>
    LOOM_DEVICE void
    store(std::uint64_t v) const
    {
        char* col = base + rows * LOOM_CX(CellLayout<K>::off[R]);
        if constexpr(ir::table_of<K>.reg[R].width == 8)
            reinterpret_cast<std::uint64_t*>(col)[row] = v;
        else
            reinterpret_cast<std::uint32_t*>(col)[row] = static_cast<std::uint32_t>(v);
    }
};

/// @brief Hide a wave-uniform pointer from the optimiser, so the column
/// addresses built on it are formed where they are used.  Without it the
/// compiler hoists every column's address out of the scheduling loop and
/// keeps them all live, as 64-bit per-lane values, for the whole kernel.
LOOM_DEVICE void
opaque_base(char*& p)
{
    asm volatile("" : "+s"(p));
}

/// @brief Run block @p B on a lane whose registers are in device memory.
template <typename K, std::size_t B>
LOOM_DEVICE void
exec_on(const MemRegs<K>& mr, std::uint32_t& pc, std::uint32_t& err)
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];
    MemRegs<K>          m   = mr;
    opaque_base(m.base);
    Regs<K>             r;
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                if constexpr(blk.use.test(R)) r.template set_raw<R>(m.template load<R>());
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
    exec_block<K, B>(r, pc, err);
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                if constexpr(blk.def.test(R) && live_out<K, B>.test(R))
                    m.template store<R>(r.template raw<R>());
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief Start thread @p tid in its row: inputs, and zero for every other
/// register live at START (load_inputs' rule).
template <typename K>
LOOM_DEVICE void
mem_load_inputs(MemRegs<K>& m, const void* const* in, std::uint64_t tid)
{
    Regs<K> r;
    load_inputs<K>(r, in, tid);
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                if constexpr(ir::table_of<K>.block[0].live_in.test(R) || ir::table_of<K>.reg[R].input)
                    m.template store<R>(r.template raw<R>());
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief Store the outputs of the halted thread @p tid from its row.
template <typename K>
LOOM_DEVICE void
mem_store_outputs(const MemRegs<K>& m, void* const* out, std::uint64_t tid)
{
    Regs<K> r;
    [&]<std::size_t... J>(std::index_sequence<J...>)
    {
        (
            [&]
            {
                constexpr std::size_t R = ir::table_of<K>.out_slot[J];
                r.template set_raw<R>(m.template load<R>());
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nout>{});
    store_outputs<K>(r, out, tid);
}
} // namespace loom::hipvm
