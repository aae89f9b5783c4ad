// -*- C++ -*-
/// @file wgr.hpp
/// @brief Regrouping inside a workgroup: the threads of a whole CU's waves
/// are sorted by their current block through LDS between run phases, so
/// each wave continues with threads at the same block, without kernel
/// boundaries and without moving registers through HBM.
///
/// Reference: the hybrid compiler's regroup (src/vm/simd.hpp:4092-4252), a
/// stable counting sort of the whole pool by program counter with halted
/// rows last (simd.hpp:2667-2745), every register column gathered and
/// scattered.  Here the pool is one workgroup of W waves (1024 threads
/// with 16 waves, every lane of a CU when a lane needs at most 128 VGPRs;
/// 512 with 8, two workgroups per CU):
///
///   sort      every lane's key is its block (HALT last); an LDS histogram
///             (one atomic per distinct key of a wave) and its prefix give
///             each thread its position in block order;
///   exchange  pc and thread id, then the registers live at the thread's
///             block, move through an LDS buffer, a fixed range of at
///             most kWgrRoundWords register words per round: each lane writes its
///             thread's live words at the thread's position and reads the
///             live words of the thread at its own index;
///   retire    the lanes at HALT (the last ones after the sort) store the
///             outputs of their finished threads; one atomic per workgroup
///             claims new threads for them from the input;
///   run       every wave runs its lanes under policy P until the policy
///             stops (successor-major: the best block has at most
///             `threshold` lanes) or the phase's dispatch budget is spent.
///
/// The loop ends, for the whole workgroup at once, when no lane holds a
/// live thread and the input is exhausted; every barrier is reached by
/// every wave of the workgroup on every iteration.
#pragma once

#include "sched.hpp"

namespace loom::hipvm
{
/// @brief Thread id of a lane that holds no thread.
inline constexpr std::uint32_t kNoTid = 0xFFFFFFFFu;
/// @brief Most state words exchanged per thread per LDS round.  A kernel
/// moves its n words in ceil(n / kWgrRoundWords) rounds of equal size (the
/// LDS buffer holds one round), so a small register file keeps a small
/// buffer and the occupancy it allows.
inline constexpr std::size_t kWgrRoundWords = 31;

/// @brief Default waves of a regrouping workgroup: 16 (one workgroup holds
/// every lane of a CU at four waves per SIMD); the other size is 8 (two
/// workgroups per CU, so one runs while the other sorts).
inline constexpr int kWgrWaves = 16;

struct WgrArgs
{
    std::uint64_t* next  = nullptr; ///< input cursor (zero at launch)
    std::uint32_t  phase = 64;      ///< dispatch budget of a run phase
};

/// @brief The 4-byte words of a thread's registers in the exchange: the
/// cells of the registers live at some block's entry or at HALT, in slot
/// order (two words for a 64-bit register).  A round
/// moves a fixed range of words, and a thread moves only the words of the
/// registers live at its key (the registers live into block k for k < nb,
/// the outputs for the key of threads at HALT): registers dead at a
/// thread's block are not moved, and the exchange code reads no other, so
/// it does not keep the whole register file live across the run phases.
/// Because a word keeps its place in every key's layout, a lane reads a
/// register's new value only after it has written the old one.
template <typename K>
struct WgrWords
{
    static constexpr std::size_t NB = ir::table_of<K>.nb;
    struct Map
    {
        std::array<std::uint16_t, 2 * ir::max_regs> reg{};
        std::array<std::uint8_t, 2 * ir::max_regs>  half{};
        std::size_t                                 n = 0;
    };
    static constexpr Map map = []
    {
        // registers live at some key only: a register that is never live at
        // a block's entry (a value used within one block) never moves
        ir::Bits any = ir::table_of<K>.outputs;
        for(std::size_t k = 0; k < NB; ++k) any = any | ir::table_of<K>.block[k].live_in;
        Map m;
        for(std::size_t r = 0; r < ir::table_of<K>.nr; ++r)
        {
            if(!any.test(r)) continue;
            m.reg[m.n]  = static_cast<std::uint16_t>(r);
            m.half[m.n] = 0;
            ++m.n;
            if(ir::table_of<K>.reg[r].width == 8)
            {
                m.reg[m.n]  = static_cast<std::uint16_t>(r);
                m.half[m.n] = 1;
                ++m.n;
            }
        }
        return m;
    }();
    static constexpr std::size_t n = map.n;

    /// word Wd belongs to a register live at key Key
    template <std::size_t Key, std::size_t Wd>
    static constexpr bool live = []
    {
        if constexpr(Wd >= n)
            return false;
        else if constexpr(Key < NB)
            return ir::table_of<K>.block[Key].live_in.test(map.reg[Wd]);
        else
            return ir::table_of<K>.outputs.test(map.reg[Wd]);
    }();
};

/// @brief Rounds of kernel @p K's exchange and words per round.
template <typename K>
struct WgrRounds
{
    static constexpr std::size_t n      = WgrWords<K>::n;
    static constexpr std::size_t rounds = n ? (n + kWgrRoundWords - 1) / kWgrRoundWords : 0;
    static constexpr std::size_t even   = rounds ? (n + rounds - 1) / rounds : 0;
    /// at least two: the pc and thread-id exchange uses the buffer's first
    /// two rows
    static constexpr std::size_t words = even < 2 ? 2 : even;
};

template <typename K, std::size_t Wd>
LOOM_DEVICE std::uint32_t
wgr_word(const Regs<K>& r)
{
    using M                 = WgrWords<K>;
    constexpr std::size_t R = M::map.reg[Wd];
    const std::uint64_t   v = r.template raw<R>();
    if constexpr(M::map.half[Wd] != 0)
        return static_cast<std::uint32_t>(v >> 32);
    else
        return static_cast<std::uint32_t>(v);
}

template <typename K, std::size_t Wd>
LOOM_DEVICE void
wgr_set_word(Regs<K>& r, std::uint32_t v)
{
    using M                 = WgrWords<K>;
    constexpr std::size_t R = M::map.reg[Wd];
    if constexpr(ir::table_of<K>.reg[R].width == 8)
    {
        const std::uint64_t old = r.template raw<R>();
        if constexpr(M::map.half[Wd] != 0)
            r.template set_raw<R>((old & 0xFFFFFFFFull) | (std::uint64_t(v) << 32));
        else
            r.template set_raw<R>((old & ~0xFFFFFFFFull) | v);
    }
    else
        r.template set_raw<R>(v);
}

/// @brief Call @p f with the compile-time key equal to @p key.
template <typename K, typename F>
LOOM_DEVICE void
wgr_on_key(std::uint32_t key, F&& f)
{
    [&]<std::size_t... Kk>(std::index_sequence<Kk...>)
    {
        ((key == Kk ? (f(std::integral_constant<std::size_t, Kk>{}), 0) : 0), ...);
    }(std::make_index_sequence<WgrWords<K>::NB + 1>{});
}

/// @brief One exchange round of words [W0, W0 + RW): a thread's
/// live words go to its position @p pos (key @p old_key), and this lane
/// takes the live words of the thread now at its index @p me (key
/// @p new_key, known from the pc word exchanged first).  Both barriers are
/// reached by every lane.
template <typename K, std::size_t W0, std::size_t RW, std::uint32_t T>
LOOM_DEVICE void
wgr_round(Regs<K>& r, std::uint32_t old_key, std::uint32_t new_key, std::uint32_t* buf,
          std::uint32_t pos, std::uint32_t me)
{
    using M = WgrWords<K>;
    wgr_on_key<K>(old_key, [&](auto kc)
    {
        constexpr std::size_t Key = decltype(kc)::value;
        [&]<std::size_t... J>(std::index_sequence<J...>)
        {
            (
                [&]
                {
                    if constexpr(M::template live<Key, W0 + J>)
                        buf[J * T + pos] = wgr_word<K, W0 + J>(r);
                }(),
                ...);
        }(std::make_index_sequence<RW>{});
    });
    __syncthreads();
    wgr_on_key<K>(new_key, [&](auto kc)
    {
        constexpr std::size_t Key = decltype(kc)::value;
        [&]<std::size_t... J>(std::index_sequence<J...>)
        {
            (
                [&]
                {
                    if constexpr(M::template live<Key, W0 + J>)
                        wgr_set_word<K, W0 + J>(r, buf[J * T + me]);
                }(),
                ...);
        }(std::make_index_sequence<RW>{});
    });
    __syncthreads();
}

/// @brief The regrouping kernel (see the file comment) with workgroups of
/// @p W waves.  A persistent grid: every workgroup loops until the input is
/// exhausted and its own threads have finished; workgroups never wait for
/// one another.  Both sizes are compiled for four waves per SIMD (128
/// VGPRs), which the 16-wave workgroup needs to be resident at all and
/// which lets two 8-wave workgroups share a CU.
template <typename K, int P, int W, bool CNT = true>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void
wgr_kernel(IoArgs io, CtlArgs c, WgrArgs g)
{
    static_assert(W == 16 || W == 8, "a regrouping workgroup has 16 or 8 waves");
    sct_table_init<K>();
    constexpr std::uint32_t T    = 64u * W;
    constexpr std::uint32_t NB   = ir::table_of<K>.nb;
    constexpr std::uint32_t KEYS = NB + 1; // blocks, then "no live thread"
    constexpr std::size_t   NR   = WgrRounds<K>::rounds;
    constexpr std::size_t   RW   = WgrRounds<K>::words;
    static_assert(KEYS <= T, "more blocks than lanes");
    __shared__ std::uint32_t hist[KEYS];
    __shared__ std::uint32_t cur[KEYS];
    __shared__ std::uint32_t s_empty, s_got, s_base_lo, s_base_hi, s_stop;
    __shared__ std::uint32_t buf[RW * T];

    const std::uint32_t me  = threadIdx.x;
    RegsOf<K, CNT>      r;
    std::uint32_t       pc  = HALT;
    std::uint32_t       tid = kNoTid;
    WaveCtx             w(c);
    const std::uint64_t budget = c.max_steps ? c.max_steps : 1;
    bool                over   = false; // the run's dispatch budget is spent
    if(me == 0) s_stop = 0;
    __syncthreads();
    for(;;)
    {
        // ---- a wave past the run's dispatch budget stops the workgroup ----
        if((w.issued + w.disp) >= budget && lane_id() == 0) s_stop = 1;
        // ---- sort by block, HALT last ----
        const std::uint32_t key = pc < NB ? pc : NB;
        if(me < KEYS)
        {
            hist[me] = 0;
            cur[me]  = 0;
        }
        __syncthreads();
        if(s_stop)
        {
            over = true;
            break;
        }
        // histogram: one LDS atomic per distinct key of a wave; a lane's
        // offset among the threads of its key is the count before its
        // wave's group plus its rank in the group
        std::uint32_t off = 0;
        for(std::uint64_t todo = ballot(true); todo;)
        {
            const std::uint32_t lead = static_cast<std::uint32_t>(__builtin_ctzll(todo));
            const std::uint32_t k    = __builtin_amdgcn_readlane(key, lead);
            const std::uint64_t m    = ballot(key == k);
            std::uint32_t       old  = 0;
            if(lane_id() == lead) old = atomicAdd(&hist[k], popc(m));
            old = __builtin_amdgcn_readlane(old, lead);
            if(key == k) off = old + rank_in(m);
            todo &= ~m;
        }
        __syncthreads();
        if(me == 0)
        {
            std::uint32_t acc = 0;
            for(std::uint32_t k = 0; k < KEYS; ++k)
            {
                cur[k] = acc;
                acc += hist[k];
            }
            s_empty = hist[NB];
        }
        __syncthreads();
        const std::uint32_t pos = cur[key] + off;
        // pc and thread id first: a lane learns which block its new thread
        // is at, and so which register words come with it
        buf[pos]     = pc;
        buf[T + pos] = tid;
        __syncthreads();
        pc  = buf[me];
        tid = buf[T + me];
        __syncthreads();
        const std::uint32_t nkey = pc < NB ? pc : NB;
        [&]<std::size_t... Rd>(std::index_sequence<Rd...>)
        { (wgr_round<K, Rd * RW, RW, T>(r, key, nkey, buf, pos, me), ...); }(std::make_index_sequence<NR>{});

        // ---- retire finished threads, claim new ones for the empty lanes ----
        if(pc == HALT && tid != kNoTid)
        {
            store_outputs<K>(r, io.out, tid);
            tid = kNoTid;
        }
        const std::uint32_t empty = s_empty;
        if(me == 0)
        {
            std::uint64_t base = 0;
            std::uint32_t got  = 0;
            if(empty)
            {
                base = atomicAdd(reinterpret_cast<unsigned long long*>(g.next),
                                 static_cast<unsigned long long>(empty));
                got  = base < c.n ? static_cast<std::uint32_t>(
                                       base + empty <= c.n ? empty : c.n - base) :
                                    0u;
            }
            s_got     = got;
            s_base_lo = static_cast<std::uint32_t>(base);
            s_base_hi = static_cast<std::uint32_t>(base >> 32);
        }
        __syncthreads();
        const std::uint32_t got   = s_got;
        const std::uint64_t base  = (std::uint64_t(s_base_hi) << 32) | s_base_lo;
        const std::uint32_t first = T - empty; // the empty lanes are the last
        if(first + got == 0) break;            // the same in every wave
        if(me >= first && me < first + got)
        {
            tid = static_cast<std::uint32_t>(base + (me - first));
            load_inputs<K>(r, io.in, tid);
            pc = 0;
        }

        // ---- run phase ----
        w.issued += w.disp;
        w.disp = 0;
        w.rem  = g.phase ? g.phase : 1;
        w.left = static_cast<std::uint32_t>(w.rem < kSegment ? w.rem : kSegment);
        w.over = false;
        run_policy<K, P>(r, pc, c.thr, 0, true, w);
    }
    flush_stats(w);
    raise_errors(w.err, over, c);
}
} // namespace loom::hipvm
