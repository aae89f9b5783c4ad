// -*- C++ -*-
/// @file stream.hpp
/// @brief The device-resident stream: a fixed set of thread slots whose
/// register files stay in HBM between steps, with exchange of finished
/// threads' outputs for new inputs done on the device.
///
/// Reference: the concurrent stream facade (src/vm/simd.hpp:2797-2903) over
/// a structure-of-arrays register file (src/vm/bve.hpp:494-693) whose rows
/// are swapped in place by `exchange_from` (bve.hpp:1208-1229).  A step runs
/// every live slot for a while (`s()`); threads that halt keep their outputs
/// in their slot; `exchange` hands those outputs out and starts new threads
/// in the same slots; `done` holds when no slot is live.
///
/// Layout of a stream of S slots (S a multiple of 64):
///   regs   one column of S elements per kernel register (4 or 8 bytes per
///          element), column r at byte offset S * (sum of the widths of the
///          registers before r)
///   pc     u32 per slot: the block the thread resumes at, kHaltPc once it
///          halted (its outputs wait in its registers), kEmptyPc for a slot
///          without a thread
///   tid    u64 per slot: sequence number of the slot's thread (set_inputs
///          numbers threads 0.., exchanges continue the numbering)
///   ring   u32 per slot: slots whose thread halted, in the order they
///          halted; a step appends to it (one atomic per wave), an exchange
///          or drain consumes a prefix.  Each slot is in it at most once, so S
///          entries suffice; positions are taken modulo S.
#pragma once

#include "sched.hpp"

namespace loom::hipvm
{
/// @brief pc of a slot whose thread halted (outputs pending).
inline constexpr std::uint32_t kHaltPc = HALT;
/// @brief pc of a slot that holds no thread.
inline constexpr std::uint32_t kEmptyPc = 0xFFFFFFFFu;

struct StreamCtl
{
    Slot64 tail; ///< ring entries appended so far (monotonic)
};

struct StreamArgs
{
    char*          regs  = nullptr;
    std::uint32_t* pc    = nullptr;
    std::uint64_t* tid   = nullptr;
    std::uint32_t* ring  = nullptr;
    StreamCtl*     ctl   = nullptr;
    std::uint64_t  slots = 0;
};

/// @brief What a swap launch does with its rows.
enum StreamSwap : std::uint32_t
{
    SWAP_INIT     = 0, ///< slots [0, S): the first m get new threads, the rest become empty
    SWAP_EXCHANGE = 1, ///< ring[head + i], i < m: outputs out, new thread in
    SWAP_DRAIN    = 2, ///< ring[head + i], i < m: outputs out, slot empty
};

/// @brief Staging columns of one swap launch (device, structure of arrays,
/// row i of each column belongs to the i-th slot swapped).
struct StreamIo
{
    const void*    in[kMaxIO]{};
    void*          out[kMaxIO]{};
    std::uint64_t* ids      = nullptr; ///< sequence numbers of the threads whose outputs leave
    std::uint64_t  head     = 0;       ///< ring position of row 0
    std::uint64_t  next_tid = 0;       ///< sequence number of the first new thread
    std::uint32_t  m        = 0;       ///< rows
    std::uint32_t  mode     = SWAP_INIT;
};

/// @brief Byte offsets of the register columns (per slot, before * S): the
/// cell layout of a register file in device memory.
template <typename K>
using StateLayout = CellLayout<K>;

/// @brief Read every register of slot @p slot into this lane.
template <typename K>
LOOM_DEVICE void
load_state(Regs<K>& r, const StreamArgs& a, std::uint64_t slot)
{
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                const char* col = a.regs + a.slots * LOOM_CX(StateLayout<K>::off[R]);
                if constexpr(ir::table_of<K>.reg[R].width == 8)
                    r.template set_raw<R>(reinterpret_cast<const std::uint64_t*>(col)[slot]);
                else
                    r.template set_raw<R>(reinterpret_cast<const std::uint32_t*>(col)[slot]);
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief Write every register of this lane to slot @p slot.
template <typename K>
LOOM_DEVICE void
store_state(const Regs<K>& r, const StreamArgs& a, std::uint64_t slot)
{
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                char* col = a.regs + a.slots * LOOM_CX(StateLayout<K>::off[R]);
                if constexpr(ir::table_of<K>.reg[R].width == 8)
                    reinterpret_cast<std::uint64_t*>(col)[slot] = r.template raw<R>();
                else
                    reinterpret_cast<std::uint32_t*>(col)[slot] =
                        static_cast<std::uint32_t>(r.template raw<R>());
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief One step of the stream: each wave owns 64 consecutive slots, loads
/// the registers of its live slots, runs them under policy @p P until at
/// most `live_thr` of them are live (or the dispatch budget of the step is
/// spent: a step budget only ends the step, the threads keep their state),
/// writes the registers back and appends the slots that halted to the ring.
template <typename K, int P>
__global__ __launch_bounds__(64) void
stream_step_kernel(StreamArgs a, CtlArgs c)
{
    sct_table_init<K>();
    const std::uint64_t slot = std::uint64_t(blockIdx.x) * 64u + lane_id();
    const std::uint32_t pc0  = a.pc[slot];
    const bool          live = pc0 < kHaltPc;
    if(!ballot(live)) return;
    Regs<K>       r;
    std::uint32_t pc = kHaltPc;
    if(live)
    {
        load_state<K>(r, a, slot);
        pc = pc0;
    }
    WaveCtx w(c);
    run_policy<K, P>(r, pc, c.thr, c.live_thr, true, w);
    if(live)
    {
        store_state<K>(r, a, slot);
        a.pc[slot] = pc;
    }
    const std::uint64_t halted = ballot(live && pc == kHaltPc);
    if(halted)
    {
        std::uint64_t base = 0;
        if(lane_id() == 0)
            base = atomicAdd(reinterpret_cast<unsigned long long*>(&a.ctl->tail.v),
                             static_cast<unsigned long long>(popc(halted)));
        base = uniform64(base);
        if(live && pc == kHaltPc) a.ring[(base + rank_in(halted)) % a.slots] = static_cast<std::uint32_t>(slot);
    }
    flush_stats(w);
    raise_errors(w.err, false, c);
}

/// @brief Start a new thread in slot @p slot from row @p row of the staged
/// input columns: registers read at START before being written are zero,
/// inputs are set (load_inputs' rule), pc = START.
template <typename K>
LOOM_DEVICE void
start_thread(const StreamArgs& a, const StreamIo& x, std::uint64_t slot, std::uint64_t row,
             std::uint64_t tid)
{
    Regs<K> r;
    load_inputs<K>(r, x.in, row);
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                constexpr auto& g = ir::table_of<K>.reg[R];
                if constexpr(ir::table_of<K>.block[0].live_in.test(R) || g.input)
                {
                    char* col = a.regs + a.slots * LOOM_CX(StateLayout<K>::off[R]);
                    if constexpr(g.width == 8)
                        reinterpret_cast<std::uint64_t*>(col)[slot] = r.template raw<R>();
                    else
                        reinterpret_cast<std::uint32_t*>(col)[slot] =
                            static_cast<std::uint32_t>(r.template raw<R>());
                }
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
    a.pc[slot]  = 0;
    a.tid[slot] = tid;
}

/// @brief Copy the outputs of the halted thread in slot @p slot to row
/// @p row of the staged output columns.
template <typename K>
LOOM_DEVICE void
emit_outputs(const StreamArgs& a, const StreamIo& x, std::uint64_t slot, std::uint64_t row)
{
    [&]<std::size_t... J>(std::index_sequence<J...>)
    {
        (
            [&]
            {
                constexpr std::size_t R = ir::table_of<K>.out_slot[J];
                using V                 = ctype<ir::table_of<K>.out_ty[J]>;
                const char* col         = a.regs + a.slots * LOOM_CX(StateLayout<K>::off[R]);
                if constexpr(ir::table_of<K>.reg[R].width == 8)
                    static_cast<V*>(x.out[J])[row] =
                        std::bit_cast<V>(reinterpret_cast<const std::uint64_t*>(col)[slot]);
                else
                {
                    const std::uint32_t v = reinterpret_cast<const std::uint32_t*>(col)[slot];
                    if constexpr(sizeof(V) == 4)
                        static_cast<V*>(x.out[J])[row] = std::bit_cast<V>(v);
                    else if constexpr(sizeof(V) == 2)
                        static_cast<V*>(x.out[J])[row] = std::bit_cast<V>(static_cast<std::uint16_t>(v));
                    else
                        static_cast<V*>(x.out[J])[row] = std::bit_cast<V>(static_cast<std::uint8_t>(v));
                }
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nout>{});
    x.ids[row] = a.tid[slot];
}

/// @brief Initialise, exchange or drain rows of a stream (StreamSwap).  One
/// lane per row; the grid covers S rows for SWAP_INIT and m rows otherwise.
template <typename K>
__global__ __launch_bounds__(256) void
stream_swap_kernel(StreamArgs a, StreamIo x)
{
    const std::uint64_t i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(x.mode == SWAP_INIT)
    {
        if(i >= a.slots) return;
        if(i < x.m)
            start_thread<K>(a, x, i, i, x.next_tid + i);
        else
            a.pc[i] = kEmptyPc;
        return;
    }
    if(i >= x.m) return;
    const std::uint64_t slot = a.ring[(x.head + i) % a.slots];
    emit_outputs<K>(a, x, slot, i);
    if(x.mode == SWAP_EXCHANGE)
        start_thread<K>(a, x, slot, i, x.next_tid + i);
    else
        a.pc[slot] = kEmptyPc;
}
} // namespace loom::hipvm
