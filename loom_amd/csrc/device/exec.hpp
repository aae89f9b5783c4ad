// -*- C++ -*-
/// @file exec.hpp
/// @brief Lowering of basic blocks to gfx950 code: a per-lane register file
/// and one inlined instruction sequence per block.
///
/// The register file is an array of 32-bit slots (registers of at most four
/// bytes) and an array of 64-bit slots, indexed only by compile-time
/// constants, so the compiler keeps every virtual register in VGPRs.  Each
/// instruction is expanded from the constant kernel table into straight-line
/// code; there is no interpretation at run time.  A block executes every one
/// of its instructions; the program counter written last decides the next
/// block (reference semantics, core.hpp:736-802, 804-829).
#pragma once

#include "ops.hpp"

#include <bit>
#include <cstdint>
#include <utility>

namespace loom::hipvm
{
using ir::halt16;
using ir::SK;

/// @brief Lane program counter value of a halted thread.
inline constexpr std::uint32_t HALT = halt16;
/// @brief "No block" result of the scheduler helpers.
inline constexpr std::uint32_t NONE = 0xFFFFFFFEu;

/// @brief Slot layout of kernel @p K's registers on a lane.
template <typename K>
struct Layout
{
    static constexpr const auto& T = ir::table_of<K>;

    struct Map
    {
        std::array<std::uint16_t, ir::max_regs> idx{};
        std::size_t                             n32 = 0, n64 = 0;
    };

    static constexpr Map map = []
    {
        Map m;
        for(std::size_t r = 0; r < ir::table_of<K>.nr; ++r)
            m.idx[r] = static_cast<std::uint16_t>(ir::table_of<K>.reg[r].width == 8 ? m.n64++ : m.n32++);
        return m;
    }();

    static constexpr std::size_t n32 = map.n32 ? map.n32 : 1;
    static constexpr std::size_t n64 = map.n64 ? map.n64 : 1;
};

/// @brief N distinct scalar cells (one base class per cell).  Unlike an
/// array, a cell can only be addressed with a compile-time index, so no
/// optimisation (e.g. merging the stores of several switch cases into one
/// store with a computed index) can force the register file out of VGPRs
/// into scratch memory.
template <std::size_t I, typename T>
struct Cell
{
    T v;
};

template <typename T, typename Seq>
struct CellsOf;

template <typename T, std::size_t... I>
struct CellsOf<T, std::index_sequence<I...>> : Cell<I, T>...
{
    template <std::size_t J>
    LOOM_DEVICE T&
    at()
    {
        return static_cast<Cell<J, T>&>(*this).v;
    }

    template <std::size_t J>
    LOOM_DEVICE const T&
    at() const
    {
        return static_cast<const Cell<J, T>&>(*this).v;
    }
};

template <typename T, std::size_t N>
using Cells = CellsOf<T, std::make_index_sequence<N>>;

template <typename K>
struct Regs
{
    using L = Layout<K>;
    Cells<std::uint32_t, L::n32> r32;
    Cells<std::uint64_t, L::n64> r64;

    template <std::size_t slot, typename V>
    LOOM_DEVICE V
    get() const
    {
        constexpr std::size_t i = L::map.idx[slot];
        if constexpr(L::T.reg[slot].width == 8)
            return std::bit_cast<V>(r64.template at<i>());
        else if constexpr(sizeof(V) == 4)
            return std::bit_cast<V>(r32.template at<i>());
        else if constexpr(sizeof(V) == 2)
            return std::bit_cast<V>(static_cast<std::uint16_t>(r32.template at<i>()));
        else
            return std::bit_cast<V>(static_cast<std::uint8_t>(r32.template at<i>()));
    }

    template <std::size_t slot, typename V>
    LOOM_DEVICE void
    set(V v)
    {
        constexpr std::size_t i = L::map.idx[slot];
        if constexpr(L::T.reg[slot].width == 8)
            r64.template at<i>() = std::bit_cast<std::uint64_t>(v);
        else if constexpr(sizeof(V) == 4)
            r32.template at<i>() = std::bit_cast<std::uint32_t>(v);
        else if constexpr(sizeof(V) == 2)
            r32.template at<i>() = std::bit_cast<std::uint16_t>(v);
        else
            r32.template at<i>() = std::bit_cast<std::uint8_t>(v);
    }

    /// @brief Raw slot bits (for spill/fill).
    template <std::size_t slot>
    LOOM_DEVICE std::uint64_t
    raw() const
    {
        constexpr std::size_t i = L::map.idx[slot];
        if constexpr(L::T.reg[slot].width == 8)
            return r64.template at<i>();
        else
            return r32.template at<i>();
    }

    template <std::size_t slot>
    LOOM_DEVICE void
    set_raw(std::uint64_t v)
    {
        constexpr std::size_t i = L::map.idx[slot];
        if constexpr(L::T.reg[slot].width == 8)
            r64.template at<i>() = v;
        else
            r32.template at<i>() = static_cast<std::uint32_t>(v);
    }
};

/// @brief Value of operand @p O (a literal or a register read).
template <typename K, std::size_t O>
LOOM_DEVICE auto
operand(const Regs<K>& r)
{
    constexpr ir::Operand o = ir::table_of<K>.opnd[O];
    using V                 = ctype<o.ty>;
    if constexpr(o.lit)
    {
        if constexpr(sizeof(V) == 8)
            return std::bit_cast<V>(static_cast<std::uint64_t>(o.bits));
        else if constexpr(sizeof(V) == 4)
            return std::bit_cast<V>(static_cast<std::uint32_t>(o.bits));
        else if constexpr(sizeof(V) == 2)
            return std::bit_cast<V>(static_cast<std::uint16_t>(o.bits));
        else
            return std::bit_cast<V>(static_cast<std::uint8_t>(o.bits));
    }
    else
        return r.template get<o.slot, V>();
}

/// @brief Bits of a table constant as type @p V.
template <typename V>
LOOM_DEVICE V
const_as(std::uint64_t bits)
{
    if constexpr(sizeof(V) == 8)
        return std::bit_cast<V>(bits);
    else if constexpr(sizeof(V) == 4)
        return std::bit_cast<V>(static_cast<std::uint32_t>(bits));
    else if constexpr(sizeof(V) == 2)
        return std::bit_cast<V>(static_cast<std::uint16_t>(bits));
    else
        return std::bit_cast<V>(static_cast<std::uint8_t>(bits));
}

/// @brief Bits of 1/d for a literal divisor d = ±2^k of the evaluation type
/// @p E whose reciprocal is a normal number; 0 for any other operand.  For
/// such a d, x / d and x · (1/d) are the same real number for every x (±0,
/// subnormals, infinities and NaN included), rounded once, so the quotient
/// is a multiplication; a correctly rounded f32 division is about eleven
/// instructions on gfx950.
template <typename K, std::size_t O, typename E>
consteval std::uint64_t
exact_reciprocal_bits()
{
    constexpr ir::Operand o = ir::table_of<K>.opnd[O];
    if constexpr(!std::is_floating_point_v<E> || !o.lit || !std::is_same_v<ctype<o.ty>, E>)
        return 0;
    else
    {
        constexpr int           mbits = sizeof(E) == 4 ? 23 : 52;
        constexpr std::uint64_t emax  = sizeof(E) == 4 ? 0xFF : 0x7FF; // all-ones exponent
        constexpr std::uint64_t b     = sizeof(E) == 4 ? (o.bits & 0xFFFFFFFFull) : o.bits;
        constexpr std::uint64_t mant  = b & ((std::uint64_t(1) << mbits) - 1);
        constexpr std::uint64_t e     = (b >> mbits) & emax;
        constexpr std::uint64_t sign  = b & ~((std::uint64_t(1) << (mbits + (sizeof(E) == 4 ? 8 : 11))) - 1);
        // d = ±2^(e - bias), bias = (emax - 1) / 2; 1/d has biased exponent
        // 2·bias - e, a normal exponent when 1 <= e <= emax - 2
        if constexpr(mant != 0 || e < 1 || e > emax - 2)
            return 0;
        else
            return sign | ((emax - 1 - e) << mbits);
    }
}

/// @brief Constant tables of `sct` (select a constant by index) and of
/// multiway `ParmJump` in LDS.
///
/// `sct(i, c0, ..., cN-1)` is c_i for 0 <= i < N and 0 otherwise
/// (operators.hpp `sct`); a multiway ParmJump on v goes to label v, or to
/// HALT when v is past its label list.  Written as a chain of compares and
/// selects either costs about two VALU instructions per entry; the path
/// tracer's shading block runs eight selections of nine constants per
/// bounce.  For a 4-byte element type (and every jump) with at least
/// kSctTableMin entries the entries go into a table in LDS, one segment per
/// instruction with a trailing 0 (HALT for a jump), and the instruction is
/// one clamp of the index (shared by every table lookup of the same index)
/// and one LDS read at a constant offset.  Every kernel that runs blocks of
/// @p K fills the table first (sct_table_init), with all its threads.
inline constexpr std::size_t kSctTableMin = 4;

template <typename K>
struct SctTable
{
    static constexpr bool
    uses(const ir::Instr& in)
    {
        if(in.kind == SK::pjump) return in.nt >= kSctTableMin;
        return in.kind == SK::compute && in.op == ir::Op::sct && ir::ty_size(in.ety) == 4 &&
               in.nc >= kSctTableMin;
    }

    /// entries of an instruction's segment before the trailing one
    static constexpr std::size_t
    entries(const ir::Instr& in)
    {
        return in.kind == SK::pjump ? in.nt : in.nc;
    }

    struct Map
    {
        std::array<std::uint16_t, ir::table_of<K>.instr.size()> off{};
        std::size_t                                            words = 0;
    };

    static constexpr Map map = []
    {
        Map m;
        for(std::size_t i = 0; i < ir::table_of<K>.ni; ++i)
        {
            const ir::Instr& in = ir::table_of<K>.instr[i];
            if(!uses(in)) continue;
            m.off[i] = static_cast<std::uint16_t>(m.words);
            m.words += entries(in) + 1u;
        }
        return m;
    }();
    static constexpr std::size_t words = map.words;

    /// the constants' bits in the element type (the low 32 bits of a 4-byte
    /// constant, as const_as reads them), each segment ending in 0; a jump's
    /// target blocks, ending in HALT
    static constexpr std::array<std::uint32_t, words ? words : 1> data = []
    {
        std::array<std::uint32_t, words ? words : 1> d{};
        for(std::size_t i = 0; i < ir::table_of<K>.ni; ++i)
        {
            const ir::Instr& in = ir::table_of<K>.instr[i];
            if(!uses(in)) continue;
            if(in.kind == SK::pjump)
            {
                for(std::size_t j = 0; j < in.nt; ++j)
                    d[map.off[i] + j] = ir::table_of<K>.target[in.t0 + j];
                d[map.off[i] + in.nt] = HALT;
                continue;
            }
            for(std::size_t j = 0; j < in.nc; ++j)
                d[map.off[i] + j] = static_cast<std::uint32_t>(ir::table_of<K>.konst[in.c0 + j]);
            d[map.off[i] + in.nc] = 0;
        }
        return d;
    }();
};

template <typename K>
__shared__ std::uint32_t sct_lds[SctTable<K>::words ? SctTable<K>::words : 1];

/// @brief Fill @p K's sct tables in LDS; every thread of the workgroup calls
/// it before any of them runs a block of @p K (it ends in a barrier).
template <typename K>
LOOM_DEVICE void
sct_table_init()
{
    if constexpr(SctTable<K>::words != 0)
    {
        if(threadIdx.x == 0)
        {
            [&]<std::size_t... W>(std::index_sequence<W...>)
            {
                ((sct_lds<K>[W] = SctTable<K>::data[W]), ...);
            }(std::make_index_sequence<SctTable<K>::words>{});
        }
        __syncthreads();
    }
}

/// @brief Run-time error flags raised by kernels.
enum ErrorFlag : std::uint32_t
{
    ERR_ASSERT_FP   = 1u << 0, ///< ops::assert_fp saw NaN/Inf
    ERR_STEP_BUDGET = 1u << 1, ///< a wave exceeded its block-step budget
    ERR_QUEUE_FULL  = 1u << 2, ///< a regroup queue overflowed (never expected)
};

/// @brief Execute instruction @p I of kernel @p K on this lane.
template <typename K, std::size_t I>
LOOM_DEVICE void
exec_instr(Regs<K>& r, std::uint32_t& pc, std::uint32_t& err)
{
    constexpr ir::Instr   in = ir::table_of<K>.instr[I];

    if constexpr(in.kind == SK::jump)
        pc = LOOM_CX(std::uint32_t(ir::table_of<K>.target[in.t0]));
    else if constexpr(in.kind == SK::pjump)
    {
        const auto v = operand<K, in.a0>(r);
        if constexpr(in.nt == 2)
            pc = v != decltype(v)(0) ? LOOM_CX(std::uint32_t(ir::table_of<K>.target[in.t0 + 1])) :
                                       LOOM_CX(std::uint32_t(ir::table_of<K>.target[in.t0]));
        else if constexpr(SctTable<K>::uses(in))
        {
            // index clamped to nt (the segment's trailing HALT)
            using UX = std::make_unsigned_t<decltype(v)>;
            const UX            u = static_cast<UX>(v);
            const std::uint32_t k = u < UX(in.nt) ? static_cast<std::uint32_t>(u) : in.nt;
            pc                    = sct_lds<K>[SctTable<K>::map.off[I] + k];
        }
        else
        {
            std::uint32_t p = HALT;
            [&]<std::size_t... J>(std::index_sequence<J...>)
            {
                ((p = (v == decltype(v)(J)) ? LOOM_CX(std::uint32_t(ir::table_of<K>.target[in.t0 + J])) : p),
                 ...);
            }(std::make_index_sequence<in.nt>{});
            pc = p;
        }
    }
    else
    {
        constexpr ir::Operand d = ir::table_of<K>.opnd[in.dst];
        using D                 = ctype<d.ty>;
        using E                 = ctype<in.ety>;
        constexpr Op op         = in.op;

        if constexpr(op == Op::mov)
            r.template set<d.slot, D>(conv<D>(operand<K, in.a0>(r)));
        else if constexpr(op == Op::assert_fp)
        {
            const E v = conv<E>(operand<K, in.a0>(r));
            if(!(v - v == E(0))) err |= ERR_ASSERT_FP;
            r.template set<d.slot, D>(conv<D>(v));
        }
        else if constexpr(op == Op::cmov)
        {
            const auto c = operand<K, in.a0>(r);
            const E    a = conv<E>(operand<K, in.a0 + 1>(r));
            const E    b = conv<E>(operand<K, in.a0 + 2>(r));
            r.template set<d.slot, D>(conv<D>(c != decltype(c)(0) ? a : b));
        }
        else if constexpr(op == Op::fma || op == Op::fms)
        {
            static_assert(std::is_floating_point_v<E>, "fma needs floating point");
            const E x = conv<E>(operand<K, in.a0>(r));
            const E y = conv<E>(operand<K, in.a0 + 1>(r));
            const E z = conv<E>(operand<K, in.a0 + 2>(r));
            r.template set<d.slot, D>(conv<D>(fma_<E>(x, y, op == Op::fma ? z : E(-z))));
        }
        else if constexpr(op == Op::sct && SctTable<K>::uses(in))
        {
            // index clamped to nc (the segment's trailing 0): negative and
            // out-of-range indices select 0, as in the select chain below
            using IX             = ctype<in.ity>;
            using UX             = std::make_unsigned_t<IX>;
            const UX          u  = static_cast<UX>(conv<IX>(operand<K, in.a0>(r)));
            const std::uint32_t k = u < UX(in.nc) ? static_cast<std::uint32_t>(u) : in.nc;
            const std::uint32_t bits = sct_lds<K>[SctTable<K>::map.off[I] + k];
            r.template set<d.slot, D>(conv<D>(std::bit_cast<E>(bits)));
        }
        else if constexpr(op == Op::sct)
        {
            using IX = ctype<in.ity>;
            const IX i = conv<IX>(operand<K, in.a0>(r));
            E       v = E(0);
            [&]<std::size_t... J>(std::index_sequence<J...>)
            {
                ((v = (i == IX(J)) ? const_as<E>(LOOM_CX(std::uint64_t(ir::table_of<K>.konst[in.c0 + J]))) : v),
                 ...);
            }(std::make_index_sequence<in.nc>{});
            r.template set<d.slot, D>(conv<D>(v));
        }
        else if constexpr(op == Op::slt)
        {
            using IX = ctype<in.ity>;
            const IX i = conv<IX>(operand<K, in.a0>(r));
            E       v = E(0);
            [&]<std::size_t... J>(std::index_sequence<J...>)
            {
                ((v = (i == IX(J)) ? conv<E>(operand<K, in.a0 + 1 + J>(r)) : v), ...);
            }(std::make_index_sequence<in.na - 1>{});
            r.template set<d.slot, D>(conv<D>(v));
        }
        else if constexpr(op == Op::pacc)
        {
            using P      = ctype<in.ity>;
            const P p    = conv<P>(operand<K, in.a0>(r));
            const E x    = conv<E>(operand<K, in.a0 + 1>(r));
            const E y    = conv<E>(operand<K, in.a0 + 2>(r));
            const bool c = arith<in.cmp, E>(x, y);
            const bool v = in.acc == Op::land ? (p != P(0)) && c : (p != P(0)) || c;
            r.template set<d.slot, D>(conv<D>(v));
        }
        else if constexpr(op == Op::neg || op == Op::lnot || op == Op::bnot ||
                          op == Op::abs || op == Op::rsqrt || op == Op::round ||
                          op == Op::ceil || op == Op::floor || op == Op::trunc ||
                          op == Op::sin || op == Op::cos || op == Op::sqrt)
        {
            const E a = conv<E>(operand<K, in.a0>(r));
            r.template set<d.slot, D>(conv<D>(arith<op, E>(a)));
        }
        else if constexpr(op == Op::div && exact_reciprocal_bits<K, in.a0 + 1, E>() != 0)
        {
            using U         = std::conditional_t<sizeof(E) == 4, std::uint32_t, std::uint64_t>;
            constexpr E rcp = std::bit_cast<E>(static_cast<U>(exact_reciprocal_bits<K, in.a0 + 1, E>()));
            const E     a   = conv<E>(operand<K, in.a0>(r));
            r.template set<d.slot, D>(conv<D>(a * rcp));
        }
        else
        {
            const E a = conv<E>(operand<K, in.a0>(r));
            const E b = conv<E>(operand<K, in.a0 + 1>(r));
            r.template set<d.slot, D>(conv<D>(arith<op, E>(a, b)));
        }
    }
}

/// @brief Execute all instructions of block @p B.
template <typename K, std::size_t B>
LOOM_DEVICE void
exec_block(Regs<K>& r, std::uint32_t& pc, std::uint32_t& err)
{
    constexpr ir::Block blk = ir::table_of<K>.block[B];
    [&]<std::size_t... J>(std::index_sequence<J...>)
    { (exec_instr<K, blk.i0 + J>(r, pc, err), ...); }(std::make_index_sequence<blk.ni>{});
    if constexpr(ir::table_of<K>.straightline) pc = HALT;
}

/// @brief Run block @p B on a lane whose registers are in VGPRs.
template <typename K, std::size_t B>
LOOM_DEVICE void
exec_on(Regs<K>& r, std::uint32_t& pc, std::uint32_t& err)
{
    exec_block<K, B>(r, pc, err);
}

/// @brief Load kernel inputs of thread @p tid into the register file and
/// zero every other register live at START.
template <typename K>
LOOM_DEVICE void
load_inputs(Regs<K>& r, const void* const* in, std::uint64_t tid)
{
    [&]<std::size_t... R>(std::index_sequence<R...>)
    {
        (
            [&]
            {
                if constexpr(ir::table_of<K>.block[0].live_in.test(R) && !ir::table_of<K>.reg[R].input)
                    r.template set_raw<R>(0);
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nr>{});
    [&]<std::size_t... J>(std::index_sequence<J...>)
    {
        (
            [&]
            {
                using V = ctype<ir::table_of<K>.in_ty[J]>;
                r.template set<ir::table_of<K>.in_slot[J], V>(static_cast<const V*>(in[J])[tid]);
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nin>{});
}

/// @brief Store kernel outputs of thread @p tid.
template <typename K>
LOOM_DEVICE void
store_outputs(const Regs<K>& r, void* const* out, std::uint64_t tid)
{
    [&]<std::size_t... J>(std::index_sequence<J...>)
    {
        (
            [&]
            {
                using V = ctype<ir::table_of<K>.out_ty[J]>;
                static_cast<V*>(out[J])[tid] = r.template get<ir::table_of<K>.out_slot[J], V>();
            }(),
            ...);
    }(std::make_index_sequence<ir::table_of<K>.nout>{});
}
} // namespace loom::hipvm
