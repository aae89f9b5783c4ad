// -*- C++ -*-
/// @file sched.hpp
/// @brief Wave64 schedulers: the predicated "compact" form and the
/// regrouping "hybrid" form.
///
/// Both keep a wave's 64 threads in VGPRs and choose, per step, one block
/// for the whole wave; EXEC is the set of lanes whose program counter equals
/// that block.  The choice follows the reference SUCCESSOR_MAJOR threaded
/// dispatch (simd.hpp:3545-3732, 3974-4005): after a block runs, the most
/// populated successor runs next while its population exceeds the threshold;
/// otherwise the most populated block of the pending set is chosen, else of
/// all blocks present in the wave (greedy groups, simd.hpp:2477-2531).
/// Populations are `popcount(ballot(pc == b))`, ties go to the lower block
/// index or the earlier successor, exactly as `major_group` (2606-2637).
///
/// The hybrid form (reference hybrid_compiler, simd.hpp:4016-4292) runs the
/// same executor inside a persistent pass kernel.  A wave takes a batch of 64
/// threads from one per-label queue (or 64 new threads), executes until the
/// best block population drops to the threshold, refills free lanes with new
/// threads (the stream exchange of simd.hpp:2878-2890 done in registers),
/// and finally spills survivors into the queue of their current label: lanes
/// of one label get consecutive rows from one atomic add (ballot + mbcnt
/// rank), and only registers live at that label are written, column by
/// column (structure of arrays).  The next pass drains those queues in full
/// 64-row batches, which is the regroup (the counting sort of
/// simd.hpp:2667-2745 without moving dead registers).
#pragma once

#include "memregs.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace loom::hipvm
{
/// @brief Waves per SIMD that the hybrid pass kernel of @p K without the
/// regroup paths is compiled to allow at least (amdgpu_waves_per_eu): 1
/// leaves the register budget to the compiler.  A kernel whose register file
/// sits just above a budget step (128 VGPRs for 4 waves, 96 for 5) can run
/// faster held under it, a few values spilled, than at one wave less; the
/// registration of such a kernel sets its hint (LOOM_HIP_WAVES_HINT).  The
/// compact kernels are left to the compiler: under the successor-major
/// policy they can need far more registers, which a hint would spill.
template <typename K>
inline constexpr int waves_hint = 1;
/// @brief The same for the pass kernel with the regroup paths, whose fill
/// and spill switches keep more registers live, so that a hint that pays
/// without them can spill too much with them (LOOM_HIP_REGROUP_WAVES_HINT).
template <typename K>
inline constexpr int regroup_waves_hint = 1;

inline constexpr int         kMaxIO    = 32;
inline constexpr std::size_t kNumStats = 13;
/// Counter copies: a wave adds into copy (workgroup index mod kStatSlots),
/// the host sums the copies, so hundreds of thousands of waves do not
/// serialise their atomics on 13 addresses.
inline constexpr std::size_t kStatSlots = 64;
inline constexpr std::size_t kStatWords = kNumStats * kStatSlots;

/// @brief Kernel input/output columns (device pointers, structure of arrays).
struct IoArgs
{
    const void* in[kMaxIO];
    void*       out[kMaxIO];
};

/// @brief Counters accumulated per run (device, u64 each).
enum Stat : int
{
    ST_ACTIVE_OPS = 0, ///< Σ popcount(EXEC)·ops(B) over block dispatches
    ST_SLOT_OPS,       ///< Σ 64·ops(B)
    ST_DISPATCHES,     ///< wave block dispatches
    ST_THREAD_STEPS,   ///< Σ popcount(EXEC): thread-block-steps
    ST_SPILL_ROWS,     ///< threads written to regroup queues
    ST_SPILL_BYTES,    ///< bytes written by regroup spills
    ST_FILL_ROWS,      ///< threads read from regroup queues
    ST_FILL_BYTES,     ///< bytes read by regroup fills
    ST_ADMITTED,       ///< threads admitted from the input
    ST_REFILLS,        ///< in-wave refill events
    ST_BATCHES,        ///< batches claimed by persistent waves
    ST_SPILL_FAILED,   ///< lanes kept because a queue was full
    ST_ACTIVE_FLOPS,   ///< Σ popcount(EXEC)·flops(B): useful floating-point work
};

/// @brief Per-run control arguments.
struct CtlArgs
{
    std::uint64_t  n         = 0;       ///< number of threads
    std::uint64_t  max_steps = 1ull << 40; ///< per-wave block dispatch budget
    std::uint32_t  thr       = 0;       ///< block-population threshold (lanes)
    std::uint32_t  live_thr  = 0;       ///< hybrid: stop when live lanes <= this
    std::uint32_t  refill    = 1;       ///< hybrid: refill free lanes in-wave
    std::uint32_t  regroup   = 1;       ///< hybrid: spill survivors to queues
    std::uint32_t* err       = nullptr; ///< error flags
    std::uint64_t* stats     = nullptr; ///< kStatSlots copies of kNumStats counters, or null
};

/// @brief Input partitions of a hybrid run: one per XCD, so that waves
/// (dispatched round-robin over the 8 XCDs) mostly claim from their own
/// counter and their own L2.
inline constexpr std::uint32_t kRanges = 8;

/// @brief Hybrid control block, resident in device memory.
/// @brief A 64-bit counter alone in its 256-byte block.  Device-scope
/// atomics from every XCD meet at one memory-side unit per cache line, so
/// counters that many waves hit must not share a line.
struct alignas(256) Slot64
{
    std::uint64_t v;
};

struct HybridCtl
{
    Slot64        next[kRanges];  ///< next unadmitted thread of each range
    Slot64        end[kRanges];   ///< end of each range
    std::uint32_t batch_next;     ///< batch claim counter of the pass
    std::uint32_t nbatches;       ///< queue batches of the pass
    std::uint32_t parity;         ///< in-queue side of the pass
    std::uint32_t live;           ///< threads queued for the pass
    std::uint32_t passes;         ///< passes run so far
    std::uint32_t done;           ///< set when no thread remains
    std::uint32_t exhausted;      ///< bit r: range r has no input left (read before claiming)
    std::uint32_t qcount[2][64];  ///< per side, per label queue fill
    std::uint32_t bprefix[65];    ///< batch prefix over labels of the pass
};

struct HybridArgs
{
    HybridCtl* ctl = nullptr;
    char*      q[2]{nullptr, nullptr}; ///< queue buffers (two sides)
    std::uint64_t qcap = 0;            ///< rows per label queue (multiple of 64)
};

// ---------------------------------------------------------------------------
// wave primitives
// ---------------------------------------------------------------------------
LOOM_DEVICE std::uint64_t
ballot(bool p)
{
    return __ballot(p);
}

LOOM_DEVICE std::uint32_t
popc(std::uint64_t m)
{
    return static_cast<std::uint32_t>(__popcll(m));
}

LOOM_DEVICE std::uint32_t
lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

/// @brief Number of set bits of @p m below this lane.
LOOM_DEVICE std::uint32_t
rank_in(std::uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<std::uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<std::uint32_t>(m), 0u));
}

LOOM_DEVICE std::uint32_t
uniform(std::uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

LOOM_DEVICE std::uint64_t
uniform64(std::uint64_t v)
{
    const std::uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<std::uint32_t>(v));
    const std::uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<std::uint32_t>(v >> 32));
    return (std::uint64_t(hi) << 32) | lo;
}

/// A flush happens at the first checkpoint with slots >= kFlushSlots.  Until
/// then a lane's ops are at most slots/64, thread steps at most slots, and
/// flops at most twice the ops.  Between checkpoints a wave makes at most
/// kSegment dispatches plus one sweep (at most 64 blocks) of at most 2^16
/// ops each, so slots grow by less than 192 · 64 · 2^16 < 2^30 and no
/// 32-bit counter wraps.
inline constexpr std::uint32_t kFlushSlots = 1u << 30;
/// Dispatches between checkpoints.  The policies that pick one block at a
/// time check after every dispatch, so a segment never passes the budget;
/// the sweep policy checks after every sweep, so the budget trips at most
/// one sweep late.
inline constexpr std::uint32_t kSegment = 128;

/// Blocks of at least this many ops count their work (thread steps, active
/// operations and flops) from a scalar popcount of the block mask, on the
/// scalar unit; smaller blocks (whose waves are short of scalar issue, not of
/// vector issue or registers) with per-lane adds.
inline constexpr std::uint32_t kScalarStepOps = 8;

/// @brief Per-wave scheduling state.  Useful work is counted per lane (a
/// VALU add under the block's EXEC mask) and summed over the wave when
/// flushed; dispatches and slots are counted in scalar registers, and the
/// budget is checked at checkpoints (after every kSegment dispatches, or
/// after every sweep of the sweep policy).
struct WaveCtx
{
    std::uint64_t  rem        = 0; ///< wave-uniform: budget left at the last checkpoint
    std::uint64_t  issued     = 0; ///< wave-uniform: dispatches up to the last checkpoint
    std::uint64_t  flushed    = 0; ///< wave-uniform: dispatches already in the run's counters
    std::uint64_t* stats      = nullptr;
    std::uint32_t  disp       = 0; ///< wave-uniform: dispatches since the last checkpoint
    std::uint32_t  left       = 0; ///< wave-uniform: dispatches of the current segment
    std::uint32_t  slots      = 0; ///< wave-uniform: Σ 64·ops(B) since the last flush
    std::uint32_t  steps      = 0; ///< wave-uniform: Σ popcount(EXEC) of large blocks since the last flush
    std::uint32_t  ops        = 0; ///< wave-uniform: Σ popcount(EXEC)·ops(B) of large blocks since the last flush
    std::uint32_t  flops      = 0; ///< wave-uniform: Σ popcount(EXEC)·flops(B) of large blocks since the last flush
    std::uint32_t  lane_steps = 0; ///< this lane: runs of small blocks
    std::uint32_t  lane_ops   = 0; ///< this lane: Σ ops(B) of the blocks it ran
    std::uint32_t  lane_flops = 0; ///< this lane: Σ flops(B) of the blocks it ran
    std::uint32_t  err        = 0; ///< this lane: error bits
    bool           over       = false;

    LOOM_DEVICE explicit WaveCtx(const CtlArgs& c) :
        rem(c.max_steps ? c.max_steps : 1), stats(c.stats),
        left(static_cast<std::uint32_t>(rem < kSegment ? rem : kSegment))
    {
    }
};

/// @brief This workgroup's copy of the counters.
LOOM_DEVICE unsigned long long*
stat_slot(std::uint64_t* stats)
{
    return reinterpret_cast<unsigned long long*>(stats + (blockIdx.x % kStatSlots) * kNumStats);
}

LOOM_DEVICE unsigned long long*
stat_slot(const CtlArgs& c)
{
    return stat_slot(c.stats);
}

/// @brief Add @p v to counter @p i of a copy (no atomic for zero).
LOOM_DEVICE void
stat_add(unsigned long long* st, std::size_t i, std::uint64_t v)
{
    if(v) atomicAdd(&st[i], static_cast<unsigned long long>(v));
}

/// @brief Sum of a per-lane u32 over the wave, whose total must fit 32
/// bits (every lane must be active).
LOOM_DEVICE std::uint32_t
wave_sum(std::uint32_t v)
{
    for(int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return uniform(v);
}

/// @brief Add the wave's counters to the run's and clear them.  Called with
/// every lane active: at the end of a kernel and at checkpoints.
LOOM_DEVICE void
flush_stats(WaveCtx& w)
{
    if(w.stats)
    {
        // a lane's ops are below slots/64 < 2^25 and its flops below twice
        // that, so the wave totals (and the scalar sums of large blocks,
        // below slots and twice slots) fit 32 bits
        const std::uint64_t act   = std::uint64_t(w.ops) + wave_sum(w.lane_ops);
        const std::uint64_t flops = std::uint64_t(w.flops) + wave_sum(w.lane_flops);
        const std::uint64_t steps = std::uint64_t(w.steps) + wave_sum(w.lane_steps);
        const std::uint64_t total = w.issued + w.disp;
        if(lane_id() == 0)
        {
            auto* st = stat_slot(w.stats);
            stat_add(st, ST_ACTIVE_OPS, act);
            stat_add(st, ST_SLOT_OPS, w.slots);
            stat_add(st, ST_DISPATCHES, total - w.flushed);
            stat_add(st, ST_THREAD_STEPS, steps);
            stat_add(st, ST_ACTIVE_FLOPS, flops);
        }
        w.flushed = total;
    }
    w.slots      = 0;
    w.steps      = 0;
    w.ops        = 0;
    w.flops      = 0;
    w.lane_steps = 0;
    w.lane_ops   = 0;
    w.lane_flops = 0;
}

/// @brief Flush when the slot count says a lane field may fill up.
LOOM_DEVICE void
flush_if_high(WaveCtx& w)
{
    if(__builtin_expect(w.slots >= kFlushSlots, 0)) flush_stats(w);
}

/// @brief Fold the dispatches since the last checkpoint into the budget (the
/// step budget trips once the wave has made max_steps dispatches), flush
/// before a counter field can overflow, and start the next segment.  Every
/// lane must be active.
LOOM_DEVICE void
checkpoint(WaveCtx& w)
{
    flush_if_high(w);
    // rem <= disp, and rem -= disp, in 32-bit halves (scalar instructions)
    const std::uint32_t hi = static_cast<std::uint32_t>(w.rem >> 32);
    const std::uint32_t lo = static_cast<std::uint32_t>(w.rem);
    w.issued += w.disp;
    if(hi == 0 && lo <= w.disp)
    {
        w.disp = 0;
        w.rem  = 0;
        w.over = true;
        return;
    }
    w.rem -= w.disp;
    w.disp = 0;
    const bool small = (w.rem >> 32) == 0 && static_cast<std::uint32_t>(w.rem) < kSegment;
    w.left           = small ? static_cast<std::uint32_t>(w.rem) : kSegment;
}

/// @brief Checkpoint at the end of a segment (after a dispatch or a sweep).
LOOM_DEVICE void
segment_check(WaveCtx& w)
{
    if(__builtin_expect(w.disp >= w.left, 0)) checkpoint(w);
}

/// @brief Step result: the next block (or rank) in the low bits, and
/// kMayHalt when the block just run has HALT among its successors (only
/// then can the number of live lanes have dropped).
inline constexpr std::uint32_t kMayHalt = 1u << 16;
inline constexpr std::uint32_t kIndex   = 0xFFFFu;

/// @brief The register file of a run that collects no statistics: small
/// blocks then skip their per-lane work counts, which cost a kernel made of
/// short blocks (Collatz's loop) about an eighth of its time.  Large blocks
/// count on the scalar unit either way.
template <typename K>
struct Uncounted : Regs<K>
{
};

template <typename RF>
inline constexpr bool counts_small_blocks = true;
template <typename K>
inline constexpr bool counts_small_blocks<Uncounted<K>> = false;

/// @brief Kernel @p K has a block below kScalarStepOps ops (only such
/// kernels are compiled a second time without per-lane counts).
template <typename K>
inline constexpr bool has_small_blocks = []
{
    for(std::size_t b = 0; b < ir::table_of<K>.nb; ++b)
        if(ir::table_of<K>.block[b].ops < kScalarStepOps) return true;
    return false;
}();

/// @brief The register file of kernel @p K's runs: counted, or not (CNT).
template <typename K, bool CNT>
using RegsOf = std::conditional_t<CNT, Regs<K>, Uncounted<K>>;

/// @brief Run block @p B on the lanes at B and count the work.
template <typename K, std::size_t B, typename RF>
LOOM_DEVICE void
run_block(RF& r, std::uint32_t& pc, WaveCtx& w)
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];
    static_assert(blk.flops <= 2u * blk.ops, "flops exceed twice the op count");
    if constexpr(blk.ops >= kScalarStepOps)
    {
        const std::uint32_t n = popc(ballot(pc == B));
        w.steps += n;
        w.ops += n * blk.ops;
        if constexpr(blk.flops != 0) w.flops += n * blk.flops;
        if(pc == B) exec_on<K, B>(r, pc, w.err);
    }
    else if(pc == B)
    {
        exec_on<K, B>(r, pc, w.err);
        if constexpr(counts_small_blocks<RF>)
        {
            w.lane_steps += 1;
            w.lane_ops += blk.ops;
            if constexpr(blk.flops != 0) w.lane_flops += blk.flops;
        }
    }
    w.slots += 64u * blk.ops;
    w.disp += 1;
}

/// @brief True when HALT is a successor of block @p B.
template <typename K, std::size_t B>
inline constexpr bool may_halt = []
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];
    for(std::size_t j = 0; j < blk.ns; ++j)
        if(ir::table_of<K>.succ[blk.s0 + j] == HALT) return true;
    return blk.ns == 0;
}();

/// @brief Mask of the blocks that have HALT among their successors.
template <typename K>
inline constexpr std::uint64_t halt_blocks = []
{
    std::uint64_t m = 0;
    [&]<std::size_t... B>(std::index_sequence<B...>)
    { ((m |= may_halt<K, B> ? 1ull << B : 0ull), ...); }(std::make_index_sequence<ir::table_of<K>.nb>{});
    return m;
}();

/// @brief Run block @p B on the lanes at B and pick the next block among its
/// successors (threaded dispatch, simd.hpp:3577-3731).
template <typename K, std::size_t B, typename RF>
LOOM_DEVICE std::uint32_t
step(RF& r, std::uint32_t& pc, std::uint64_t& pending, std::uint32_t thr, WaveCtx& w)
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];

    run_block<K, B>(r, pc, w);
    pending &= ~(1ull << B);

    if constexpr(blk.ns == 0)
        return NONE;
    else if constexpr(blk.ns == 1)
    {
        constexpr std::uint32_t s = ir::table_of<K>.succ[blk.s0];
        if constexpr(s == HALT)
            return NONE;
        else
            return popc(ballot(pc == s)) > thr ? s : NONE;
    }
    else
    {
        std::uint32_t best = NONE, bp = 0;
        [&]<std::size_t... J>(std::index_sequence<J...>)
        {
            (
                [&]
                {
                    constexpr std::uint32_t s = ir::table_of<K>.succ[blk.s0 + J];
                    const std::uint32_t     p = popc(ballot(pc == s));
                    if constexpr(s != HALT)
                        if(p) pending |= 1ull << s;
                    if(p > bp)
                    {
                        bp   = p;
                        best = s;
                    }
                }(),
                ...);
        }(std::make_index_sequence<blk.ns>{});
        return (bp > thr && best != HALT) ? best : NONE;
    }
}

#define LOOM_CASE(b)                                                        \
    case b:                                                                 \
        if constexpr(b < ir::table_of<K>.nb)                                \
            return step<K, (b < ir::table_of<K>.nb ? b : 0)>(r, pc, pending, thr, w); \
        else                                                                \
            return NONE;

/// @brief Jump to the step of block @p cur (a wave-uniform value).
template <typename K, typename RF>
LOOM_DEVICE std::uint32_t
dispatch(std::uint32_t cur, RF& r, std::uint32_t& pc, std::uint64_t& pending,
         std::uint32_t thr, WaveCtx& w)
{
    switch(cur)
    {
        LOOM_CASE(0) LOOM_CASE(1) LOOM_CASE(2) LOOM_CASE(3) LOOM_CASE(4)
        LOOM_CASE(5) LOOM_CASE(6) LOOM_CASE(7) LOOM_CASE(8) LOOM_CASE(9)
        LOOM_CASE(10) LOOM_CASE(11) LOOM_CASE(12) LOOM_CASE(13) LOOM_CASE(14)
        LOOM_CASE(15) LOOM_CASE(16) LOOM_CASE(17) LOOM_CASE(18) LOOM_CASE(19)
        LOOM_CASE(20) LOOM_CASE(21) LOOM_CASE(22) LOOM_CASE(23) LOOM_CASE(24)
        LOOM_CASE(25) LOOM_CASE(26) LOOM_CASE(27) LOOM_CASE(28) LOOM_CASE(29)
        LOOM_CASE(30) LOOM_CASE(31) LOOM_CASE(32) LOOM_CASE(33) LOOM_CASE(34)
        LOOM_CASE(35) LOOM_CASE(36) LOOM_CASE(37) LOOM_CASE(38) LOOM_CASE(39)
        LOOM_CASE(40) LOOM_CASE(41) LOOM_CASE(42) LOOM_CASE(43) LOOM_CASE(44)
        LOOM_CASE(45) LOOM_CASE(46) LOOM_CASE(47) LOOM_CASE(48) LOOM_CASE(49)
        LOOM_CASE(50) LOOM_CASE(51) LOOM_CASE(52) LOOM_CASE(53) LOOM_CASE(54)
        LOOM_CASE(55) LOOM_CASE(56) LOOM_CASE(57) LOOM_CASE(58) LOOM_CASE(59)
        LOOM_CASE(60) LOOM_CASE(61) LOOM_CASE(62) LOOM_CASE(63)
    default: return NONE;
    }
}
#undef LOOM_CASE

/// @brief Most populated block of the bit set @p set (ascending, strict >).
LOOM_DEVICE std::uint32_t
major_of_set(std::uint64_t set, std::uint32_t pc, std::uint32_t& pops)
{
    std::uint32_t best = NONE, bp = 0;
    while(set)
    {
        const std::uint32_t b = static_cast<std::uint32_t>(__builtin_ctzll(set));
        set &= set - 1;
        const std::uint32_t p = popc(ballot(pc == b));
        if(p > bp)
        {
            bp   = p;
            best = b;
        }
    }
    pops = bp;
    return best;
}

/// @brief Most populated block among all live lanes (greedy groups).
LOOM_DEVICE std::uint32_t
major_present(std::uint32_t pc, std::uint32_t& pops)
{
    std::uint64_t live = ballot(pc != HALT);
    std::uint32_t best = NONE, bp = 0;
    while(live)
    {
        const int           l = __builtin_ctzll(live);
        const std::uint32_t b = uniform(__builtin_amdgcn_readlane(pc, l));
        const std::uint64_t m = ballot(pc == b);
        live &= ~m;
        const std::uint32_t p = popc(m);
        if(p > bp || (p == bp && b < best))
        {
            bp   = p;
            best = b;
        }
    }
    pops = bp;
    return best;
}

/// @brief Execute the wave while the chosen block has more than @p thr lanes
/// and more than @p live_thr lanes are live.  With @p force the first choice
/// runs regardless of the thresholds, which guarantees progress for every
/// claimed batch.  Returns when no block has more than @p thr lanes (thr = 0:
/// when every lane halted) or when at most @p live_thr lanes remain live.
template <typename K, typename RF>
LOOM_DEVICE void
run_to_threshold(RF& r, std::uint32_t& pc, std::uint32_t thr, std::uint32_t live_thr,
                 bool force, WaveCtx& w)
{
    std::uint64_t pending = 0;
    std::uint32_t pops    = 0;
    std::uint32_t cur     = major_present(pc, pops);
    if(!(pops > thr || (force && pops > 0))) return;
    if(!force && live_thr && popc(ballot(pc != HALT)) <= live_thr) return;
    for(;;)
    {
        while(cur != NONE)
        {
            const std::uint32_t ran = cur;
            cur = dispatch<K>(cur, r, pc, pending, thr, w);
            segment_check(w);
            if(w.over) return;
            if(live_thr && ((LOOM_CX(halt_blocks<K>) >> ran) & 1u) &&
               popc(ballot(pc != HALT)) <= live_thr)
                return;
        }
        pops = 0;
        if(pending) cur = major_of_set(pending, pc, pops);
        if(!pops) cur = major_present(pc, pops);
        if(pops <= thr) return;
    }
}

/// @brief Reconvergence policies.
enum Policy : int
{
    /// reference SUCCESSOR_MAJOR threaded dispatch (simd.hpp:3974-4005)
    POLICY_SUCCESSOR_MAJOR = 0,
    /// lowest reverse-post-order block first; a loop header waits while any
    /// lane is inside its loop body (reconvergence at the loop head)
    POLICY_TOPO = 1,
    /// predicated sweeps: every block in reverse post-order, each run under
    /// the mask of its lanes and skipped when it has none; lanes that jump
    /// forward run in the same sweep, lanes that jump back wait for the next
    /// (so loop heads reconverge as under POLICY_TOPO) — no per-block choice,
    /// popcount or present set, which suits kernels of small blocks
    POLICY_SWEEP = 2,
};

inline constexpr int kNumPolicies = 3;

/// @brief Ranks of the topological policy: block order[R] has rank R, so
/// the present set is kept as a mask of ranks and the next block is the
/// lowest eligible rank (one find-first-set).
template <typename K>
struct TopoRanks
{
    static constexpr std::size_t nb = ir::table_of<K>.nb;

    static constexpr auto rank = []
    {
        std::array<std::uint32_t, 64> a{};
        for(std::size_t i = 0; i < nb; ++i) a[ir::table_of<K>.order[i]] = static_cast<std::uint32_t>(i);
        return a;
    }();

    /// body masks of the loop heads, in rank space
    static constexpr auto body = []
    {
        std::array<std::uint64_t, 64> a{};
        for(std::size_t h = 0; h < ir::table_of<K>.nhead; ++h)
            for(std::size_t b = 0; b < nb; ++b)
                if((ir::table_of<K>.body[h] >> b) & 1u) a[h] |= 1ull << rank[b];
        return a;
    }();
};

/// @brief Rank mask of the blocks that hold live lanes.
template <typename K>
LOOM_DEVICE std::uint64_t
present_ranks(std::uint32_t pc)
{
    std::uint64_t live = ballot(pc != HALT), pres = 0;
    while(live)
    {
        const int           l = __builtin_ctzll(live);
        const std::uint32_t b = uniform(__builtin_amdgcn_readlane(pc, l));
        live &= ~ballot(pc == b);
        switch(b)
        {
#define LOOM_RC(i)                                                                    \
    case i:                                                                           \
        if constexpr(i < ir::table_of<K>.nb) pres |= 1ull << LOOM_CX(TopoRanks<K>::rank[i]); \
        break;
            LOOM_RC(0) LOOM_RC(1) LOOM_RC(2) LOOM_RC(3) LOOM_RC(4) LOOM_RC(5) LOOM_RC(6)
            LOOM_RC(7) LOOM_RC(8) LOOM_RC(9) LOOM_RC(10) LOOM_RC(11) LOOM_RC(12) LOOM_RC(13)
            LOOM_RC(14) LOOM_RC(15) LOOM_RC(16) LOOM_RC(17) LOOM_RC(18) LOOM_RC(19)
            LOOM_RC(20) LOOM_RC(21) LOOM_RC(22) LOOM_RC(23) LOOM_RC(24) LOOM_RC(25)
            LOOM_RC(26) LOOM_RC(27) LOOM_RC(28) LOOM_RC(29) LOOM_RC(30) LOOM_RC(31)
            LOOM_RC(32) LOOM_RC(33) LOOM_RC(34) LOOM_RC(35) LOOM_RC(36) LOOM_RC(37)
            LOOM_RC(38) LOOM_RC(39) LOOM_RC(40) LOOM_RC(41) LOOM_RC(42) LOOM_RC(43)
            LOOM_RC(44) LOOM_RC(45) LOOM_RC(46) LOOM_RC(47) LOOM_RC(48) LOOM_RC(49)
            LOOM_RC(50) LOOM_RC(51) LOOM_RC(52) LOOM_RC(53) LOOM_RC(54) LOOM_RC(55)
            LOOM_RC(56) LOOM_RC(57) LOOM_RC(58) LOOM_RC(59) LOOM_RC(60) LOOM_RC(61)
            LOOM_RC(62) LOOM_RC(63)
#undef LOOM_RC
        default: break;
        }
    }
    return pres;
}

/// @brief Next rank under POLICY_TOPO among the ranks of @p present.
template <typename K>
LOOM_DEVICE std::uint32_t
topo_pick(std::uint64_t present)
{
    if(!present) return NONE;
    std::uint64_t elig = present;
    [&]<std::size_t... H>(std::index_sequence<H...>)
    {
        ((elig &= (present & LOOM_CX(TopoRanks<K>::body[H]))
                      ? ~(1ull << LOOM_CX(TopoRanks<K>::rank[ir::table_of<K>.head[H]]))
                      : ~0ull),
         ...);
    }(std::make_index_sequence<ir::table_of<K>.nhead>{});
    if(!elig) elig = present;
    return static_cast<std::uint32_t>(__builtin_ctzll(elig));
}

/// @brief POLICY_TOPO step: run the block of rank @p R, update the present
/// set from its successors, pick the next rank.
template <typename K, std::size_t R, typename RF>
LOOM_DEVICE std::uint32_t
step_topo(RF& r, std::uint32_t& pc, std::uint64_t& present, WaveCtx& w)
{
    constexpr std::size_t   B   = ir::table_of<K>.order[R];
    constexpr ir::Block     blk = ir::table_of<K>.block[B];
    constexpr std::uint32_t mh  = may_halt<K, B> ? kMayHalt : 0u;
    run_block<K, B>(r, pc, w);
    present &= ~(1ull << R);
    [&]<std::size_t... J>(std::index_sequence<J...>)
    {
        (
            [&]
            {
                constexpr std::uint32_t s = ir::table_of<K>.succ[blk.s0 + J];
                if constexpr(s != HALT)
                    if(ballot(pc == s)) present |= 1ull << LOOM_CX(TopoRanks<K>::rank[s]);
            }(),
            ...);
    }(std::make_index_sequence<blk.ns>{});
    const std::uint32_t next = topo_pick<K>(present);
    return next == NONE ? NONE : (next | mh);
}

#define LOOM_TCASE(b)                                                                     \
    case b:                                                                               \
        if constexpr(b < ir::table_of<K>.nb)                                              \
            return step_topo<K, (b < ir::table_of<K>.nb ? b : 0)>(r, pc, present, w);     \
        else                                                                              \
            return NONE;

template <typename K, typename RF>
LOOM_DEVICE std::uint32_t
dispatch_topo(std::uint32_t cur, RF& r, std::uint32_t& pc, std::uint64_t& present,
              WaveCtx& w)
{
    switch(cur)
    {
        LOOM_TCASE(0) LOOM_TCASE(1) LOOM_TCASE(2) LOOM_TCASE(3) LOOM_TCASE(4)
        LOOM_TCASE(5) LOOM_TCASE(6) LOOM_TCASE(7) LOOM_TCASE(8) LOOM_TCASE(9)
        LOOM_TCASE(10) LOOM_TCASE(11) LOOM_TCASE(12) LOOM_TCASE(13) LOOM_TCASE(14)
        LOOM_TCASE(15) LOOM_TCASE(16) LOOM_TCASE(17) LOOM_TCASE(18) LOOM_TCASE(19)
        LOOM_TCASE(20) LOOM_TCASE(21) LOOM_TCASE(22) LOOM_TCASE(23) LOOM_TCASE(24)
        LOOM_TCASE(25) LOOM_TCASE(26) LOOM_TCASE(27) LOOM_TCASE(28) LOOM_TCASE(29)
        LOOM_TCASE(30) LOOM_TCASE(31) LOOM_TCASE(32) LOOM_TCASE(33) LOOM_TCASE(34)
        LOOM_TCASE(35) LOOM_TCASE(36) LOOM_TCASE(37) LOOM_TCASE(38) LOOM_TCASE(39)
        LOOM_TCASE(40) LOOM_TCASE(41) LOOM_TCASE(42) LOOM_TCASE(43) LOOM_TCASE(44)
        LOOM_TCASE(45) LOOM_TCASE(46) LOOM_TCASE(47) LOOM_TCASE(48) LOOM_TCASE(49)
        LOOM_TCASE(50) LOOM_TCASE(51) LOOM_TCASE(52) LOOM_TCASE(53) LOOM_TCASE(54)
        LOOM_TCASE(55) LOOM_TCASE(56) LOOM_TCASE(57) LOOM_TCASE(58) LOOM_TCASE(59)
        LOOM_TCASE(60) LOOM_TCASE(61) LOOM_TCASE(62) LOOM_TCASE(63)
    default: return NONE;
    }
}
#undef LOOM_TCASE

/// @brief POLICY_TOPO runner: execute until every lane halted or at most
/// @p live_thr lanes are live (the first step runs regardless with @p force).
template <typename K, typename RF>
LOOM_DEVICE void
run_topo(RF& r, std::uint32_t& pc, std::uint32_t live_thr, bool force, WaveCtx& w)
{
    std::uint64_t present = present_ranks<K>(pc);
    std::uint32_t cur     = topo_pick<K>(present);
    if(cur == NONE) return;
    if(!force && live_thr && popc(ballot(pc != HALT)) <= live_thr) return;
    for(;;)
    {
        const std::uint32_t res = dispatch_topo<K>(cur, r, pc, present, w);
        segment_check(w);
        if(w.over || res == NONE) return;
        if(live_thr && (res & kMayHalt) && popc(ballot(pc != HALT)) <= live_thr) return;
        cur = res & kIndex;
    }
}

/// @brief POLICY_SWEEP runner: sweeps until every lane halted or, after a
/// sweep, at most @p live_thr lanes are live (one sweep runs regardless when
/// @p force is set).
template <typename K, typename RF>
LOOM_DEVICE void
run_sweep(RF& r, std::uint32_t& pc, std::uint32_t live_thr, bool force, WaveCtx& w)
{
    if(!ballot(pc != HALT)) return;
    if(!force && live_thr && popc(ballot(pc != HALT)) <= live_thr) return;
    for(;;)
    {
        [&]<std::size_t... R>(std::index_sequence<R...>)
        {
            (
                [&]
                {
                    constexpr std::size_t B = ir::table_of<K>.order[R];
                    if(ballot(pc == B)) run_block<K, B>(r, pc, w);
                }(),
                ...);
        }(std::make_index_sequence<ir::table_of<K>.nb>{});
        segment_check(w);
        if(w.over) return;
        const std::uint64_t live = ballot(pc != HALT);
        if(!live || (live_thr && popc(live) <= live_thr)) return;
    }
}

/// @brief Run the wave under policy @p P (see run_to_threshold / run_topo /
/// run_sweep).
template <typename K, int P, typename RF>
LOOM_DEVICE void
run_policy(RF& r, std::uint32_t& pc, std::uint32_t thr, std::uint32_t live_thr, bool force,
           WaveCtx& w)
{
    if constexpr(P == POLICY_TOPO)
        run_topo<K>(r, pc, live_thr, force, w);
    else if constexpr(P == POLICY_SWEEP)
        run_sweep<K>(r, pc, live_thr, force, w);
    else
        run_to_threshold<K>(r, pc, thr, live_thr, force, w);
}

LOOM_DEVICE void
raise_errors(std::uint32_t lane_err, bool over, const CtlArgs& c)
{
    std::uint32_t e = lane_err;
    if(over) e |= ERR_STEP_BUDGET;
    // OR-reduce via ballots of each bit (few bits)
    std::uint32_t we = 0;
    for(std::uint32_t b = 0; b < 3; ++b)
        if(ballot((e >> b) & 1u)) we |= 1u << b;
    if(we && c.err && lane_id() == 0) atomicOr(c.err, we);
}

// ---------------------------------------------------------------------------
// compact form
// ---------------------------------------------------------------------------
/// @brief Predicated wave64 kernel: each wave runs 64 consecutive threads to
/// completion (reference compact_compiler, simd.hpp:3742-4014).
template <typename K, int P, bool CNT = true>
__global__ __launch_bounds__(256) void
compact_kernel(IoArgs io, CtlArgs c)
{
    sct_table_init<K>();
    const std::uint64_t tid   = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool          valid = tid < c.n;
    RegsOf<K, CNT>      r;
    std::uint32_t       pc = HALT;
    if(valid)
    {
        load_inputs<K>(r, io.in, tid);
        pc = 0;
    }
    WaveCtx w(c);
    run_policy<K, P>(r, pc, 0, 0, true, w);
    if(valid) store_outputs<K>(r, io.out, tid);
    flush_stats(w);
    raise_errors(w.err, w.over, c);
}

/// @brief The register file of a compact run in device memory: column r of
/// `rows` cells at base + rows * CellLayout<K>::off[r].
struct MemArgs
{
    char*         base = nullptr;
    std::uint64_t rows = 0;
};

/// @brief The compact kernel with its register file in device memory: as
/// compact_kernel, but every block run loads the registers it reads from the
/// thread's row and stores back those it writes that stay live, so VGPRs hold
/// one block's working set (sweep policy).
template <typename K, int P>
__global__ __launch_bounds__(256) void
compact_mem_kernel(IoArgs io, CtlArgs c, MemArgs a)
{
    sct_table_init<K>();
    const std::uint64_t tid   = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool          valid = tid < c.n;
    MemRegs<K>          r{a.base, a.rows, static_cast<std::uint32_t>(tid)};
    std::uint32_t       pc = HALT;
    if(valid)
    {
        mem_load_inputs<K>(r, io.in, tid);
        pc = 0;
    }
    WaveCtx w(c);
    run_policy<K, P>(r, pc, 0, 0, true, w);
    if(valid) mem_store_outputs<K>(r, io.out, tid);
    flush_stats(w);
    raise_errors(w.err, w.over, c);
}

// ---------------------------------------------------------------------------
// hybrid form: queue layout
// ---------------------------------------------------------------------------
/// @brief Byte layout of the regroup queues of kernel @p K: label b owns a
/// region of `qcap` rows; its columns are the thread index (u32) followed by
/// the registers live at b, each a contiguous array of `qcap` elements.
template <typename K>
struct QueueLayout
{
    static constexpr const auto& T = ir::table_of<K>;

    /// row bytes of each label and their prefix
    static constexpr auto rows = []
    {
        std::array<std::uint32_t, 65> pre{};
        for(std::size_t b = 0; b < ir::table_of<K>.nb; ++b)
        {
            std::uint32_t w = 4;
            for(std::size_t s = 0; s < ir::table_of<K>.nr; ++s)
                if(ir::table_of<K>.block[b].live_in.test(s)) w += ir::table_of<K>.reg[s].width;
            pre[b + 1] = pre[b] + w;
        }
        return pre;
    }();

    /// total bytes per queue row set (multiply by qcap for one side)
    static constexpr std::uint32_t total = rows[ir::table_of<K>.nb];

    /// column prefix (bytes per row before slot @p s) inside label @p b
    static constexpr std::uint32_t
    col(std::size_t b, std::size_t s)
    {
        std::uint32_t w = 4;
        for(std::size_t k = 0; k < s; ++k)
            if(ir::table_of<K>.block[b].live_in.test(k)) w += ir::table_of<K>.reg[k].width;
        return w;
    }

    static constexpr std::uint32_t
    live_bytes(std::size_t b)
    {
        return rows[b + 1] - rows[b];
    }
};

/// @brief Hide @p v from the optimiser (an empty asm that "modifies" it), so
/// address arithmetic built on it stays at its use.  Without it the
/// compiler hoists the column addresses of every label's queue rows out of
/// the pass loop and keeps them live in spilled scalar registers for the
/// whole kernel, which costs vector registers and occupancy for code that
/// runs only when threads are spilled or filled.
LOOM_DEVICE void
opaque(std::uint64_t& v)
{
    asm volatile("" : "+s"(v));
}

/// @brief Write this lane's thread @p tid (at label @p B) to row @p row.
template <typename K, std::size_t B>
LOOM_DEVICE void
spill_row(const Regs<K>& r, std::uint32_t tid, char* q, std::uint64_t qcap, std::uint32_t row)
{
    opaque(qcap);
    using QL                = QueueLayout<K>;
    char* base              = q + qcap * LOOM_CX(QL::rows[B]);
    reinterpret_cast<std::uint32_t*>(base)[row] = tid;
    [&]<std::size_t... S>(std::index_sequence<S...>)
    {
        (
            [&]
            {
                if constexpr(ir::table_of<K>.block[B].live_in.test(S))
                {
                    char* col = base + qcap * LOOM_CX(QL::col(B, S));
                    if constexpr(ir::table_of<K>.reg[S].width == 8)
                        reinterpret_cast<std::uint64_t*>(col)[row] = r.template raw<S>();
                    else
                        reinterpret_cast<std::uint32_t*>(col)[row] =
                            static_cast<std::uint32_t>(r.template raw<S>());
                }
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
}

/// @brief Read thread state at label @p B from row @p row into this lane.
template <typename K, std::size_t B>
LOOM_DEVICE std::uint32_t
fill_row(Regs<K>& r, const char* q, std::uint64_t qcap, std::uint32_t row)
{
    opaque(qcap);
    using QL                = QueueLayout<K>;
    const char* base        = q + qcap * LOOM_CX(QL::rows[B]);
    const std::uint32_t tid = reinterpret_cast<const std::uint32_t*>(base)[row];
    [&]<std::size_t... S>(std::index_sequence<S...>)
    {
        (
            [&]
            {
                if constexpr(ir::table_of<K>.block[B].live_in.test(S))
                {
                    const char* col = base + qcap * LOOM_CX(QL::col(B, S));
                    if constexpr(ir::table_of<K>.reg[S].width == 8)
                        r.template set_raw<S>(reinterpret_cast<const std::uint64_t*>(col)[row]);
                    else
                        r.template set_raw<S>(reinterpret_cast<const std::uint32_t*>(col)[row]);
                }
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
    return tid;
}

#define LOOM_LABEL_SWITCH(EXPR)                                            \
    switch(label)                                                          \
    {                                                                      \
        LOOM_LC(0) LOOM_LC(1) LOOM_LC(2) LOOM_LC(3) LOOM_LC(4) LOOM_LC(5)  \
        LOOM_LC(6) LOOM_LC(7) LOOM_LC(8) LOOM_LC(9) LOOM_LC(10) LOOM_LC(11) \
        LOOM_LC(12) LOOM_LC(13) LOOM_LC(14) LOOM_LC(15) LOOM_LC(16)        \
        LOOM_LC(17) LOOM_LC(18) LOOM_LC(19) LOOM_LC(20) LOOM_LC(21)        \
        LOOM_LC(22) LOOM_LC(23) LOOM_LC(24) LOOM_LC(25) LOOM_LC(26)        \
        LOOM_LC(27) LOOM_LC(28) LOOM_LC(29) LOOM_LC(30) LOOM_LC(31)        \
        LOOM_LC(32) LOOM_LC(33) LOOM_LC(34) LOOM_LC(35) LOOM_LC(36)        \
        LOOM_LC(37) LOOM_LC(38) LOOM_LC(39) LOOM_LC(40) LOOM_LC(41)        \
        LOOM_LC(42) LOOM_LC(43) LOOM_LC(44) LOOM_LC(45) LOOM_LC(46)        \
        LOOM_LC(47) LOOM_LC(48) LOOM_LC(49) LOOM_LC(50) LOOM_LC(51)        \
        LOOM_LC(52) LOOM_LC(53) LOOM_LC(54) LOOM_LC(55) LOOM_LC(56)        \
        LOOM_LC(57) LOOM_LC(58) LOOM_LC(59) LOOM_LC(60) LOOM_LC(61)        \
        LOOM_LC(62) LOOM_LC(63)                                            \
    default: break;                                                        \
    }

template <typename K>
LOOM_DEVICE void
spill_label(std::uint32_t label, const Regs<K>& r, std::uint32_t tid, char* q,
            std::uint64_t qcap, std::uint32_t row)
{
#define LOOM_LC(b)                                                              \
    case b:                                                                     \
        if constexpr(b < ir::table_of<K>.nb)                                    \
            spill_row<K, (b < ir::table_of<K>.nb ? b : 0)>(r, tid, q, qcap, row); \
        break;
    LOOM_LABEL_SWITCH(0)
#undef LOOM_LC
}

/// @brief Row bytes of label @p label (a switch over compile-time values, so
/// no table is read from memory).
template <typename K>
LOOM_DEVICE std::uint32_t
label_bytes(std::uint32_t label)
{
#define LOOM_LC(b)                                                              \
    case b:                                                                     \
        if constexpr(b < ir::table_of<K>.nb)                                    \
            return LOOM_CX(QueueLayout<K>::live_bytes(b < ir::table_of<K>.nb ? b : 0)); \
        break;
    LOOM_LABEL_SWITCH(0)
#undef LOOM_LC
    return 0;
}

template <typename K>
LOOM_DEVICE std::uint32_t
fill_label(std::uint32_t label, Regs<K>& r, const char* q, std::uint64_t qcap,
           std::uint32_t row)
{
    std::uint32_t tid = 0;
#define LOOM_LC(b)                                                                  \
    case b:                                                                         \
        if constexpr(b < ir::table_of<K>.nb)                                        \
            tid = fill_row<K, (b < ir::table_of<K>.nb ? b : 0)>(r, q, qcap, row);   \
        break;
    LOOM_LABEL_SWITCH(0)
#undef LOOM_LC
    return tid;
}

// ---------------------------------------------------------------------------
// hybrid form: pass kernel
// ---------------------------------------------------------------------------
/// @brief A wave's private reservoir of unadmitted thread indices.  Refills
/// draw from it without atomics; it is topped up from the shared input
/// ranges in guided chunks (large while much input remains, one batch near
/// the end), preferring the range of the wave's XCD.  Exhausted ranges are
/// remembered so that they are not polled again.
struct Reservoir
{
    std::uint64_t base      = 0;
    std::uint32_t count     = 0;
    std::uint32_t exhausted = 0; ///< bit r: range r has no input left
    std::uint32_t chunk     = 0; ///< next claim size (0: not sized yet)
};

/// @brief Guided claim size: about half a fair share of what a range still
/// holds among the waves that draw from it, one batch at least, 512 at most.
LOOM_DEVICE std::uint32_t
guided_chunk(std::uint64_t left)
{
    const std::uint64_t share = gridDim.x / kRanges ? gridDim.x / kRanges : 1;
    const std::uint64_t c     = left / (2 * share);
    return static_cast<std::uint32_t>(c < 64 ? 64 : (c > 512 ? 512 : c));
}

/// @brief Take up to @p want consecutive thread indices (wave-uniform).
LOOM_DEVICE std::uint32_t
take_input(HybridCtl* ctl, std::uint32_t home, Reservoir& res, std::uint32_t want,
           std::uint64_t& base)
{
    if(res.count == 0)
    {
        if(res.chunk == 0)
        {
            const std::uint64_t nx = ctl->next[home].v, en = ctl->end[home].v;
            res.chunk              = guided_chunk(uniform64(nx < en ? en - nx : 0));
        }
        std::uint64_t b = 0, e = 0;
        std::uint32_t got = 0, exh = res.exhausted;
        if(lane_id() == 0)
        {
            // ranges another wave found empty: a plain load instead of one
            // more serialised atomic per wave and range at the end of input
            exh |= __hip_atomic_load(&ctl->exhausted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for(std::uint32_t i = 0; i < kRanges && !got; ++i)
            {
                const std::uint32_t r = (home + i) % kRanges;
                if((exh >> r) & 1u) continue;
                e                     = ctl->end[r].v;
                const std::uint64_t k = atomicAdd(
                    reinterpret_cast<unsigned long long*>(&ctl->next[r].v),
                    static_cast<unsigned long long>(res.chunk));
                if(k < e)
                {
                    b   = k;
                    got = static_cast<std::uint32_t>(e - k < res.chunk ? e - k : res.chunk);
                }
                else
                {
                    exh |= 1u << r;
                    __hip_atomic_fetch_or(&ctl->exhausted, 1u << r, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        res.base      = uniform64(b);
        res.count     = uniform(got);
        res.exhausted = uniform(exh);
        // guided chunking on what the range still holds
        res.chunk = guided_chunk(uniform64(e) - (res.base + res.count));
    }
    const std::uint32_t g = res.count < want ? res.count : want;
    base                  = res.base;
    res.base += g;
    res.count -= g;
    return g;
}

/// @brief One hybrid pass.  Launched with one wave per workgroup and a grid
/// sized to the resident wave count; every wave loops claiming batches until
/// the pass has none left, so each wave exits after finite work.  Waves
/// beyond the pass's useful work exit at once.
///
/// @p RG compiles the regroup paths (queue fills and spills, a switch over
/// every label that reads or writes the registers live there).  Without
/// them the register file is not held live across those paths and the
/// kernel needs far fewer VGPRs for a large register file (the reference's
/// pt2_const: 239 -> 138, two -> three waves per SIMD), so runs that do not
/// regroup launch the RG = false kernel, in which no queue is ever read.
template <typename K, int P, bool RG, bool CNT = true>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RG ? regroup_waves_hint<K> : waves_hint<K>))) void
hybrid_pass_kernel(IoArgs io, CtlArgs c, HybridArgs h)
{
    sct_table_init<K>();
    HybridCtl* ctl            = h.ctl;
    const std::uint32_t par   = uniform(ctl->parity);
    const std::uint32_t nbq   = uniform(ctl->nbatches);
    const std::uint32_t lane  = lane_id();
    const std::uint32_t home  = blockIdx.x % kRanges;

    // useful waves: queue batches plus input batches left
    std::uint64_t left = 0;
    for(std::uint32_t i = 0; i < kRanges; ++i)
    {
        const std::uint64_t nx = ctl->next[i].v, e = ctl->end[i].v;
        left += nx < e ? (e - nx + 63) / 64 : 0;
    }
    left = uniform64(left);
    if(std::uint64_t(blockIdx.x) >= nbq + left) return;
    // drain pass: the input is exhausted and every queued batch has a wave
    // of its own, so regrouping again cannot fill waves better than this
    // pass does; survivors finish in place (no pass per remaining block)
    const bool drain = left == 0 && nbq <= gridDim.x;

    [[maybe_unused]] const char*    qin  = h.q[par];
    [[maybe_unused]] char*          qout = h.q[par ^ 1u];
    [[maybe_unused]] std::uint32_t* qcnt = ctl->qcount[par ^ 1u];

    WaveCtx w(c);
    std::uint64_t spill_rows = 0, spill_bytes = 0, fill_rows = 0, fill_bytes = 0,
                  admitted = 0, refills = 0, batches = 0, spill_failed = 0;
    Reservoir res;

    RegsOf<K, CNT> r;
    std::uint32_t  pc  = HALT;
    std::uint32_t  tid = 0;
    bool           has = false;

    for(;;)
    {
        flush_if_high(w);
        // ---- claim a batch (the wave is empty here) ----
        pc              = HALT;
        has             = false;
        std::uint32_t b = nbq;
        if(RG && lane == 0 && nbq) b = atomicAdd(&ctl->batch_next, 1u);
        b = uniform(b);
        if(RG && b < nbq)
        {
            if constexpr(RG)
            {
                // label of batch b: last label whose prefix is <= b
                std::uint32_t label = 0;
                while(label + 1 < 64 && ctl->bprefix[label + 1] <= b) ++label;
                label                    = uniform(label);
                const std::uint32_t cnt  = uniform(ctl->qcount[par][label]);
                const std::uint32_t row0 = (b - uniform(ctl->bprefix[label])) * 64u;
                const std::uint32_t rows = cnt - row0 < 64u ? cnt - row0 : 64u;
                if(lane < rows)
                {
                    tid = fill_label<K>(label, r, qin, h.qcap, row0 + lane);
                    pc  = label;
                    has = true;
                }
                fill_rows += rows;
                fill_bytes += std::uint64_t(rows) * label_bytes<K>(label);
            }
        }
        else
        {
            std::uint64_t       base = 0;
            const std::uint32_t got  = take_input(ctl, home, res, 64u, base);
            if(!got) break;
            if(lane < got)
            {
                const std::uint64_t t = base + lane;
                load_inputs<K>(r, io.in, t);
                tid = static_cast<std::uint32_t>(t);
                pc  = 0;
                has = true;
            }
            admitted += got;
        }
        ++batches;

        // ---- execute, refill, spill ----
        // One call site of the block code (it is inlined: a second site
        // would double the kernel's instruction footprint).  `finish` runs
        // the remaining lanes to HALT in place.
        bool force = true, finish = false;
        for(;;)
        {
            run_policy<K, P>(r, pc, finish ? 0u : c.thr, finish ? 0u : c.live_thr, force || finish,
                             w);
            force = false;
            if(w.over) break;
            if(has && pc == HALT)
            {
                store_outputs<K>(r, io.out, tid);
                has = false;
            }
            const std::uint64_t live = ballot(has);
            if(!live || finish) break;
            if(c.refill && ~live)
            {
                const std::uint64_t empty = ~live;
                std::uint64_t       base  = 0;
                const std::uint32_t got   = take_input(ctl, home, res, popc(empty), base);
                if(got)
                {
                    const std::uint32_t k = rank_in(empty);
                    if(!has && k < got)
                    {
                        const std::uint64_t t = base + k;
                        load_inputs<K>(r, io.in, t);
                        tid = static_cast<std::uint32_t>(t);
                        pc  = 0;
                        has = true;
                    }
                    admitted += got;
                    ++refills;
                    force = true;
                    continue;
                }
            }
            if constexpr(RG)
            {
                if(c.regroup && !drain)
                {
                    // spill survivors into the queues of their labels
                    std::uint64_t todo = live;
                    while(todo)
                    {
                        const int           l     = __builtin_ctzll(todo);
                        const std::uint32_t label = uniform(__builtin_amdgcn_readlane(pc, l));
                        const std::uint64_t m     = ballot(has && pc == label);
                        todo &= ~m;
                        const std::uint32_t cntm = popc(m);
                        std::uint32_t       base = 0;
                        if(lane == 0) base = atomicAdd(&qcnt[label], cntm);
                        base = uniform(base);
                        const std::uint32_t row = base + rank_in(m);
                        if(has && pc == label && row < h.qcap)
                        {
                            spill_label<K>(label, r, tid, qout, h.qcap, row);
                            has = false;
                            pc  = HALT;
                        }
                        const std::uint32_t ok =
                            base >= h.qcap ? 0u :
                                             (h.qcap - base < cntm ? std::uint32_t(h.qcap - base) : cntm);
                        spill_rows += ok;
                        spill_bytes += std::uint64_t(ok) * label_bytes<K>(label);
                        spill_failed += cntm - ok;
                    }
                    if(!ballot(has)) break;
                }
            }
            // no regroup (or a full queue): finish the remaining lanes in place
            finish = true;
        }
        if(w.over) break;
    }
    flush_stats(w);
    if(c.stats && lane == 0)
    {
        auto* st = stat_slot(c);
        stat_add(st, ST_SPILL_ROWS, spill_rows);
        stat_add(st, ST_SPILL_BYTES, spill_bytes);
        stat_add(st, ST_FILL_ROWS, fill_rows);
        stat_add(st, ST_FILL_BYTES, fill_bytes);
        stat_add(st, ST_ADMITTED, admitted);
        stat_add(st, ST_REFILLS, refills);
        stat_add(st, ST_BATCHES, batches);
        stat_add(st, ST_SPILL_FAILED, spill_failed);
    }
    raise_errors(w.err, w.over, c);
}

/// @brief Between passes: turn the queues written by the finished pass into
/// the batches of the next one (clamped to capacity), clear the other side
/// and raise `done` when no thread is queued and no input remains.  With
/// @p first set it initialises the control block.  One wave, lane 0 works.
inline __global__ __launch_bounds__(64) void
hybrid_setup_kernel(HybridCtl* ctl, std::uint32_t nb, std::uint64_t qcap,
                    std::uint64_t n, std::uint32_t first)
{
    if(threadIdx.x != 0) return;
    if(first)
    {
        for(std::uint32_t r = 0; r < kRanges; ++r)
        {
            ctl->next[r].v = n * r / kRanges;
            ctl->end[r].v  = n * (r + 1) / kRanges;
        }
        ctl->batch_next = 0;
        ctl->nbatches   = 0;
        ctl->parity     = 0;
        ctl->live       = 0;
        ctl->passes     = 0;
        ctl->done       = n == 0;
        ctl->exhausted  = 0;
        for(std::uint32_t r = 0; r < kRanges; ++r)
            if(ctl->next[r].v >= ctl->end[r].v) ctl->exhausted |= 1u << r;
        for(int side = 0; side < 2; ++side)
            for(int l = 0; l < 64; ++l) ctl->qcount[side][l] = 0;
        for(int l = 0; l < 65; ++l) ctl->bprefix[l] = 0;
        return;
    }
    const std::uint32_t out  = ctl->parity ^ 1u;
    std::uint32_t       live = 0, pre = 0;
    for(std::uint32_t l = 0; l < 64; ++l)
    {
        std::uint32_t cnt = l < nb ? ctl->qcount[out][l] : 0u;
        if(cnt > qcap) cnt = static_cast<std::uint32_t>(qcap);
        ctl->qcount[out][l] = cnt;
        ctl->bprefix[l]     = pre;
        pre += (cnt + 63u) / 64u;
        live += cnt;
    }
    ctl->bprefix[64] = pre;
    for(std::uint32_t l = 0; l < 64; ++l) ctl->qcount[out ^ 1u][l] = 0;
    bool input_left = false;
    std::uint32_t exhausted = 0;
    for(std::uint32_t r = 0; r < kRanges; ++r)
    {
        if(ctl->next[r].v > ctl->end[r].v) ctl->next[r].v = ctl->end[r].v;
        input_left |= ctl->next[r].v < ctl->end[r].v;
        if(ctl->next[r].v >= ctl->end[r].v) exhausted |= 1u << r;
    }
    ctl->exhausted  = exhausted;
    ctl->parity     = out;
    ctl->nbatches   = pre;
    ctl->batch_next = 0;
    ctl->live       = live;
    ctl->passes += 1;
    ctl->done = live == 0 && !input_left;
}
} // namespace loom::hipvm
