// -*- C++ -*-
/// @file lower.hpp
/// @brief Compile-time lowering of a `kernel_blank` into a flat kernel table.
///
/// The reference computes block structure with deep type-list recursion
/// (core.hpp:686-1031), which needs ~30 GB of compiler memory for the path
/// tracer (README.md:245-249).  Here the statement types are walked once to
/// collect plain descriptors; everything else is a constant-evaluated
/// function over flat arrays:
///
///  1. label optimisation: leading labels alias START, trailing labels alias
///     HALT, adjacent labels merge (core.hpp:849-867, resolved transitively);
///  2. blocks split at labels; START = 0, stated labels 1..n, HALT = -1
///     (core.hpp:977-993);
///  3. fall-through jumps appended to blocks that do not end in a jump
///     (core.hpp:804-829); successors = targets of each block's last
///     instruction (core.hpp:376-393);
///  4. registers inferred by name; one name may carry several types of one
///     size, sharing one slot (core.hpp:1000-1009);
///  5. per-block use/def sets and live-in sets (a backward data-flow fixpoint;
///     HALT's live-in set is the outputs) — the lowering spills only live
///     registers when it regroups threads.
///
/// The table is the single description shared by the host facade and the
/// gfx950 code generator.  Malformed kernels are rejected at compile time by
/// calling the non-constexpr `lowering_error`, which names the problem in the
/// compiler diagnostic.
#pragma once

#include <vm/core.hpp>
#include <vm/literal.hpp>
#include <vm/operators.hpp>

#include <array>
#include <cstdint>
#include <functional>
#include <string_view>
#include <type_traits>

namespace loom::vm::ir
{
/// @brief Called during constant evaluation to reject a kernel; the call makes
/// the evaluation non-constant and the diagnostic shows @p what.
void lowering_error(const char* what);

/// @brief Scalar types a register or literal may have.
enum class Ty : std::uint8_t
{
    u8,
    u16,
    u32,
    u64,
    i8,
    i16,
    i32,
    i64,
    f32,
    f64,
    none
};

template <typename T>
consteval Ty
ty_of()
{
    using U = std::remove_cv_t<T>;
    if constexpr(std::is_same_v<U, float>)
        return Ty::f32;
    else if constexpr(std::is_same_v<U, double>)
        return Ty::f64;
    else if constexpr(std::is_integral_v<U> && !std::is_same_v<U, bool>)
    {
        constexpr bool s = std::is_signed_v<U>;
        if constexpr(sizeof(U) == 1) return s ? Ty::i8 : Ty::u8;
        if constexpr(sizeof(U) == 2) return s ? Ty::i16 : Ty::u16;
        if constexpr(sizeof(U) == 4) return s ? Ty::i32 : Ty::u32;
        if constexpr(sizeof(U) == 8) return s ? Ty::i64 : Ty::u64;
    }
    else
        static_assert(sizeof(U) == 0, "unsupported register type");
}

constexpr int
ty_size(Ty t) noexcept
{
    switch(t)
    {
    case Ty::u8:
    case Ty::i8: return 1;
    case Ty::u16:
    case Ty::i16: return 2;
    case Ty::u32:
    case Ty::i32:
    case Ty::f32: return 4;
    case Ty::u64:
    case Ty::i64:
    case Ty::f64: return 8;
    default: return 0;
    }
}

constexpr bool
ty_float(Ty t) noexcept
{
    return t == Ty::f32 || t == Ty::f64;
}

constexpr bool
ty_signed(Ty t) noexcept
{
    return t == Ty::i8 || t == Ty::i16 || t == Ty::i32 || t == Ty::i64;
}

/// @brief Operation codes (reference operators.hpp + std functors).
enum class Op : std::uint8_t
{
    mov,
    add,
    sub,
    mul,
    div,
    mod,
    neg,
    eq,
    ne,
    lt,
    gt,
    le,
    ge,
    land,
    lor,
    lnot,
    band,
    bor,
    bxor,
    bnot,
    shl,
    shr,
    cmov,
    abs,
    fma,
    fms,
    rsqrt,
    sct,
    slt,
    pacc,
    assert_fp,
    min,
    max,
    round,
    ceil,
    floor,
    trunc,
    sin,
    cos,
    sqrt,
    copysign,
    none
};

/// @brief Statement kinds.
enum class SK : std::uint8_t
{
    compute,
    label,
    jump,
    pjump
};

/// @brief Block index used for HALT in jump tables.
inline constexpr std::uint16_t halt16 = 0xFFFF;

// ---------------------------------------------------------------------------
// op traits: functor type -> op code and operand type
// ---------------------------------------------------------------------------
template <typename OpT>
struct op_traits
{
    static constexpr bool known = false;
};

#define LOOM_OP_T(FN, CODE)                     \
    template <typename T>                       \
    struct op_traits<FN<T>>                     \
    {                                           \
        static constexpr bool known = true;     \
        static constexpr Op   op    = Op::CODE; \
        using type                  = T;        \
    };
LOOM_OP_T(std::plus, add)
LOOM_OP_T(std::minus, sub)
LOOM_OP_T(std::multiplies, mul)
LOOM_OP_T(std::divides, div)
LOOM_OP_T(std::modulus, mod)
LOOM_OP_T(std::negate, neg)
LOOM_OP_T(std::equal_to, eq)
LOOM_OP_T(std::not_equal_to, ne)
LOOM_OP_T(std::less, lt)
LOOM_OP_T(std::greater, gt)
LOOM_OP_T(std::less_equal, le)
LOOM_OP_T(std::greater_equal, ge)
LOOM_OP_T(std::logical_and, land)
LOOM_OP_T(std::logical_or, lor)
LOOM_OP_T(std::logical_not, lnot)
LOOM_OP_T(std::bit_and, band)
LOOM_OP_T(std::bit_or, bor)
LOOM_OP_T(std::bit_xor, bxor)
LOOM_OP_T(std::bit_not, bnot)
LOOM_OP_T(ops::lshift, shl)
LOOM_OP_T(ops::rshift, shr)
LOOM_OP_T(ops::abs, abs)
LOOM_OP_T(ops::rsqrt, rsqrt)
LOOM_OP_T(ops::min, min)
LOOM_OP_T(ops::max, max)
LOOM_OP_T(ops::round, round)
LOOM_OP_T(ops::ceil, ceil)
LOOM_OP_T(ops::floor, floor)
LOOM_OP_T(ops::trunc, trunc)
LOOM_OP_T(ops::sin, sin)
LOOM_OP_T(ops::cos, cos)
LOOM_OP_T(ops::sqrt, sqrt)
LOOM_OP_T(ops::copysign, copysign)
#undef LOOM_OP_T

#define LOOM_OP_U(FN, CODE)                     \
    template <>                                 \
    struct op_traits<FN>                        \
    {                                           \
        static constexpr bool known = true;     \
        static constexpr Op   op    = Op::CODE; \
        using type                  = void;     \
    };
LOOM_OP_U(std::identity, mov)
LOOM_OP_U(ops::cmov, cmov)
LOOM_OP_U(ops::fma, fma)
LOOM_OP_U(ops::fms, fms)
LOOM_OP_U(ops::assert_fp, assert_fp)
#undef LOOM_OP_U

template <typename T, typename I, typename L>
struct op_traits<ops::sct<T, I, L>>
{
    static constexpr bool known = true;
    static constexpr Op   op    = Op::sct;
    using type                  = T;
    using index                 = I;
};

template <typename T, typename I>
struct op_traits<ops::slt<T, I>>
{
    static constexpr bool known = true;
    static constexpr Op   op    = Op::slt;
    using type                  = T;
    using index                 = I;
};

template <typename T,
          typename I,
          template <typename> typename Acc,
          template <typename> typename Cmp>
struct op_traits<ops::pacc<T, I, Acc, Cmp>>
{
    static constexpr bool known = true;
    static constexpr Op   op    = Op::pacc;
    using type                  = T;
    using index                 = I;
    static constexpr Op acc     = op_traits<Acc<I>>::op;
    static constexpr Op cmp     = op_traits<Cmp<T>>::op;
};

// ---------------------------------------------------------------------------
// raw collection of statements from types
// ---------------------------------------------------------------------------
/// @brief A register reference or a literal.
struct Operand
{
    std::string_view name{};
    Ty               ty   = Ty::none;
    bool             lit  = false;
    std::uint64_t    bits = 0;   ///< literal bit pattern
    std::uint16_t    slot = 0;   ///< register slot (filled by the lowering)
};

/// @brief A collected statement.
struct Stmt
{
    SK            kind  = SK::compute;
    Op            op    = Op::none;
    Ty            ety   = Ty::none; ///< evaluation type of the operation
    Ty            ity   = Ty::none; ///< index type (sct/slt) or predicate type (pacc)
    Op            acc   = Op::none; ///< pacc accumulator
    Op            cmp   = Op::none; ///< pacc comparison
    std::uint16_t dst   = 0;        ///< operand index of the destination
    std::uint16_t a0    = 0, na = 0; ///< operand range of the arguments
    std::uint16_t l0    = 0, nl = 0; ///< label range (jumps) / label name (labels)
    std::uint16_t c0    = 0, nc = 0; ///< constant range (sct)
};

struct Counts
{
    std::size_t stmts = 0, opnds = 0, labels = 0, consts = 0;

    constexpr Counts&
    operator+=(const Counts& o) noexcept
    {
        stmts += o.stmts;
        opnds += o.opnds;
        labels += o.labels;
        consts += o.consts;
        return *this;
    }
};

template <typename S>
struct stmt_info
{
    static constexpr Counts counts{};
};

template <typename... Ts>
struct stmt_info<type_sequence<Ts...>>
{
    static constexpr Counts counts = []
    {
        Counts c{};
        (c += ... += stmt_info<Ts>::counts);
        return c;
    }();
};

template <typename OpT, typename Dst, typename... Args>
struct stmt_info<Compute<OpT, Dst, Args...>>
{
    static constexpr Counts counts{1, 1 + sizeof...(Args), 0, [] {
                                       if constexpr(op_traits<OpT>::op == Op::sct)
                                           return OpT::list::size;
                                       else
                                           return std::size_t{0};
                                   }()};
};

template <typename N>
struct stmt_info<Label<N>>
{
    static constexpr Counts counts{1, 0, 1, 0};
};

template <typename L>
struct stmt_info<Jump<L>>
{
    static constexpr Counts counts{1, 0, 1, 0};
};

template <typename V, typename... Ls>
struct stmt_info<ParmJump<V, type_sequence<Ls...>>>
{
    static constexpr Counts counts{1, 1, sizeof...(Ls), 0};
};

template <std::size_t NS, std::size_t NO, std::size_t NL, std::size_t NC>
struct Raw
{
    std::array<Stmt, NS>             stmt{};
    std::array<Operand, NO>          opnd{};
    std::array<std::string_view, NL> label{};
    std::array<std::uint64_t, NC>    konst{};
    std::size_t                      ns = 0, no = 0, nl = 0, nc = 0;
};

template <typename V>
constexpr Operand
make_operand()
{
    using T = vartype<V>;
    Operand o;
    o.name       = V::name::view;
    o.ty         = ty_of<T>();
    const auto p = literal::parse<T>(o.name);
    if(p)
    {
        o.lit  = true;
        o.bits = literal::bits_of<T>(*p);
    }
    return o;
}

template <typename Cs, typename T>
constexpr std::uint64_t
const_bits()
{
    constexpr auto p = literal::parse<T>(Cs::view);
    if constexpr(!p.has_value())
        static_assert(sizeof(Cs) == 0, "sct constant is not a number");
    return literal::bits_of<T>(*p);
}

template <typename S>
struct collector;

template <typename... Ts>
struct collector<type_sequence<Ts...>>
{
    template <typename R>
    static constexpr void
    run(R& r)
    {
        (collector<Ts>::run(r), ...);
    }
};

template <typename... Ts>
struct collector<macro_inv<Ts...>>
{
    template <typename R>
    static constexpr void
    run(R&)
    {
    }
};

template <typename OpT, typename Dst, typename... Args>
struct collector<Compute<OpT, Dst, Args...>>
{
    using tr = op_traits<OpT>;
    static_assert(tr::known, "unknown Compute operation");

    template <typename R>
    static constexpr void
    run(R& r)
    {
        Stmt s;
        s.kind = SK::compute;
        s.op   = tr::op;
        if constexpr(!std::is_void_v<typename tr::type>)
            s.ety = ty_of<typename tr::type>();
        else if constexpr(tr::op == Op::cmov)
        {
            static_assert(sizeof...(Args) == 3, "cmov takes 3 arguments");
            using A = std::tuple_element_t<1, std::tuple<Args...>>;
            using B = std::tuple_element_t<2, std::tuple<Args...>>;
            static_assert(std::is_same_v<vartype<A>, vartype<B>>,
                          "cmov operands must have one type");
            s.ety = ty_of<vartype<A>>();
        }
        else
        {
            using A = std::tuple_element_t<0, std::tuple<Args...>>;
            s.ety   = ty_of<vartype<A>>();
        }
        if constexpr(tr::op == Op::sct || tr::op == Op::slt ||
                     tr::op == Op::pacc)
            s.ity = ty_of<typename tr::index>();
        if constexpr(tr::op == Op::pacc)
        {
            s.acc = tr::acc;
            s.cmp = tr::cmp;
        }
        s.dst             = static_cast<std::uint16_t>(r.no);
        r.opnd[r.no++]    = make_operand<Dst>();
        s.a0              = static_cast<std::uint16_t>(r.no);
        s.na              = sizeof...(Args);
        ((r.opnd[r.no++] = make_operand<Args>()), ...);
        if constexpr(tr::op == Op::sct)
        {
            s.c0 = static_cast<std::uint16_t>(r.nc);
            s.nc = static_cast<std::uint16_t>(OpT::list::size);
            fill_consts<typename tr::type>(r, typename OpT::list{});
        }
        r.stmt[r.ns++] = s;
    }

    template <typename T, typename R, typename... Cs>
    static constexpr void
    fill_consts(R& r, type_sequence<Cs...>)
    {
        ((r.konst[r.nc++] = const_bits<Cs, T>()), ...);
    }
};

template <typename N>
struct collector<Label<N>>
{
    template <typename R>
    static constexpr void
    run(R& r)
    {
        Stmt s;
        s.kind           = SK::label;
        s.l0             = static_cast<std::uint16_t>(r.nl);
        s.nl             = 1;
        r.label[r.nl++]  = N::view;
        r.stmt[r.ns++]   = s;
    }
};

template <typename N>
struct collector<Jump<Label<N>>>
{
    template <typename R>
    static constexpr void
    run(R& r)
    {
        Stmt s;
        s.kind          = SK::jump;
        s.l0            = static_cast<std::uint16_t>(r.nl);
        s.nl            = 1;
        r.label[r.nl++] = N::view;
        r.stmt[r.ns++]  = s;
    }
};

template <typename V, typename... Ns>
struct collector<ParmJump<V, type_sequence<Label<Ns>...>>>
{
    static_assert(sizeof...(Ns) >= 2, "ParmJump needs at least two labels");

    template <typename R>
    static constexpr void
    run(R& r)
    {
        Stmt s;
        s.kind         = SK::pjump;
        s.a0           = static_cast<std::uint16_t>(r.no);
        s.na           = 1;
        r.opnd[r.no++] = make_operand<V>();
        s.ety          = r.opnd[s.a0].ty;
        s.l0           = static_cast<std::uint16_t>(r.nl);
        s.nl           = sizeof...(Ns);
        ((r.label[r.nl++] = Ns::view), ...);
        r.stmt[r.ns++] = s;
    }
};

// ---------------------------------------------------------------------------
// the kernel table
// ---------------------------------------------------------------------------
inline constexpr std::size_t max_regs = 256;

/// @brief Register bit set.
struct Bits
{
    std::array<std::uint64_t, max_regs / 64> w{};

    constexpr void
    set(std::size_t i) noexcept
    {
        w[i / 64] |= std::uint64_t{1} << (i % 64);
    }

    constexpr bool
    test(std::size_t i) const noexcept
    {
        return (w[i / 64] >> (i % 64)) & 1u;
    }

    constexpr bool
    operator==(const Bits&) const noexcept = default;
};

constexpr Bits
operator|(Bits a, const Bits& b) noexcept
{
    for(std::size_t i = 0; i < a.w.size(); ++i) a.w[i] |= b.w[i];
    return a;
}

constexpr Bits
minus(Bits a, const Bits& b) noexcept
{
    for(std::size_t i = 0; i < a.w.size(); ++i) a.w[i] &= ~b.w[i];
    return a;
}

/// @brief One executable instruction of a block.
struct Instr
{
    SK            kind = SK::compute;
    Op            op   = Op::none;
    Ty            ety  = Ty::none;
    Ty            ity  = Ty::none;
    Op            acc  = Op::none;
    Op            cmp  = Op::none;
    std::uint16_t dst  = 0;      ///< operand index
    std::uint16_t a0 = 0, na = 0; ///< operand range
    std::uint16_t t0 = 0, nt = 0; ///< target range (block indices, halt16)
    std::uint16_t c0 = 0, nc = 0; ///< constant range
};

/// @brief A register slot.
struct Reg
{
    std::string_view name{};
    Ty               ty       = Ty::none; ///< first type seen
    std::uint8_t     width    = 0;        ///< bytes
    std::uint8_t     ntypes   = 0;        ///< >1: a union register
    bool             written  = false;
    bool             input    = false;
    bool             output   = false;
};

/// @brief A basic block.
struct Block
{
    std::uint16_t i0 = 0, ni = 0;  ///< instruction range
    std::uint16_t s0 = 0, ns = 0;  ///< successor range (block index or halt16)
    std::uint16_t ops = 0;         ///< #Compute + 1: the cost unit of a step
    std::uint16_t flops = 0;       ///< floating-point operations (fma = 2)
    Bits          use{}, def{}, live_in{};
};

template <std::size_t NS, std::size_t NO, std::size_t NL, std::size_t NC>
struct Table
{
    std::string_view kname{};
    std::uint8_t     index_bytes = 0;
    bool             straightline = false;

    std::array<Instr, NS + NL + 1>     instr{};
    std::array<Operand, NO + 1>        opnd{};
    std::array<std::uint16_t, NL + 1>  target{};
    std::array<std::uint64_t, NC + 1>  konst{};
    std::array<Block, NL + 2>          block{};
    std::array<std::uint16_t, NS + NL + 2> succ{};
    std::array<Reg, NO + 1>            reg{};
    std::array<std::uint16_t, NO + 1>  in_slot{}, out_slot{};
    std::array<Ty, NO + 1>             in_ty{}, out_ty{};
    std::array<std::string_view, NL + 2> block_label{};

    std::size_t ni = 0, no = 0, nt = 0, nc = 0, nb = 0, nsucc = 0, nr = 0,
                nin = 0, nout = 0;
    std::size_t state_size = 0;  ///< Σ register widths (core.hpp:1027-1030)
    /// scheduling order of the topological policy: blocks outside loops, then
    /// loop blocks, each group in reverse post-order of a DFS from START
    std::array<std::uint8_t, 64> order{};
    std::array<std::uint8_t, 64> rank{};
    /// loop headers (targets of DFS back edges) and their loop bodies
    std::size_t                   nhead = 0;
    std::array<std::uint8_t, 64>  head{};
    std::array<std::uint64_t, 64> body{}; ///< block mask, header excluded
    Bits        outputs{};
    std::uint64_t hash = 0;
};

constexpr std::uint64_t
fnv(std::uint64_t h, std::uint64_t v) noexcept
{
    for(int i = 0; i < 8; ++i)
    {
        h ^= (v >> (8 * i)) & 0xFF;
        h *= 0x100000001B3ull;
    }
    return h;
}

constexpr std::uint64_t
fnv(std::uint64_t h, std::string_view s) noexcept
{
    for(char c: s)
    {
        h ^= static_cast<unsigned char>(c);
        h *= 0x100000001B3ull;
    }
    return fnv(h, s.size());
}

/// @brief Collect and lower kernel @p K.
template <Kernel K>
consteval auto
lower()
{
    constexpr Counts cs  = stmt_info<typename K::statements>::counts;
    constexpr Counts ci  = stmt_info<typename K::input_vars>::counts;
    constexpr std::size_t NIn  = K::input_vars::size;
    constexpr std::size_t NOut = K::output_vars::size;
    constexpr std::size_t NS = cs.stmts, NO = cs.opnds + NIn + NOut,
                          NL = cs.labels, NC = cs.consts;
    (void)ci;

    Raw<NS + 1, NO + 1, NL + 1, NC + 1> raw{};
    collector<typename K::statements>::run(raw);

    Table<NS, NO, NL, NC> t{};
    t.kname       = K::nameview;
    t.index_bytes = sizeof(typename K::index_t);

    // kernel inputs/outputs become operands too (for register inference)
    const std::size_t in0 = raw.no;
    [&]<typename... Vs>(type_sequence<Vs...>)
    { ((raw.opnd[raw.no++] = make_operand<Vs>()), ...); }(
        typename K::input_vars{});
    const std::size_t out0 = raw.no;
    [&]<typename... Vs>(type_sequence<Vs...>)
    { ((raw.opnd[raw.no++] = make_operand<Vs>()), ...); }(
        typename K::output_vars{});

    // ---- 1. label optimisation ------------------------------------------
    // alias[i]: for a label statement i, the label name it resolves to.
    // Labels before the first non-label statement alias START, labels after
    // the last non-label statement alias HALT, runs of adjacent labels alias
    // the last label of the run.
    std::size_t first = raw.ns, last = 0;
    bool        any   = false;
    for(std::size_t i = 0; i < raw.ns; ++i)
        if(raw.stmt[i].kind != SK::label)
        {
            if(!any) first = i;
            last = i;
            any  = true;
        }
    if(!any) lowering_error("kernel has no instructions");

    std::array<std::string_view, NL + 1> alias_from{}, alias_to{};
    std::size_t                          nalias = 0;
    auto label_name = [&](std::size_t i) { return raw.label[raw.stmt[i].l0]; };
    for(std::size_t i = 0; i < raw.ns; ++i)
    {
        if(raw.stmt[i].kind != SK::label) continue;
        std::string_view to{};
        if(i < first)
            to = "START";
        else if(i > last)
            to = "HALT";
        else
        {
            std::size_t j = i;
            while(j + 1 < raw.ns && raw.stmt[j + 1].kind == SK::label) ++j;
            if(j == i) continue; // a kept label
            to = label_name(j);
        }
        alias_from[nalias] = label_name(i);
        alias_to[nalias]   = to;
        ++nalias;
    }
    auto resolve = [&](std::string_view n)
    {
        for(std::size_t k = 0; k < nalias; ++k)
            if(alias_from[k] == n) return alias_to[k];
        return n;
    };

    // ---- 2. blocks -----------------------------------------------------------
    // kept labels in order; block b starts at kept label b-1 (block 0 = START)
    std::array<std::size_t, NL + 2> bstart{}; // statement index of block start
    std::size_t                     nb = 1;
    bstart[0]                          = first;
    t.block_label[0]                   = "START";
    for(std::size_t i = first; i <= last; ++i)
    {
        if(raw.stmt[i].kind != SK::label) continue;
        if(i + 1 <= last && raw.stmt[i + 1].kind == SK::label) continue;
        const std::string_view n = label_name(i);
        if(n == "START" || n == "HALT")
            lowering_error("START and HALT cannot be stated labels");
        for(std::size_t b = 0; b < nb; ++b)
            if(t.block_label[b] == n) lowering_error("duplicate label");
        t.block_label[nb] = n;
        bstart[nb++]      = i + 1;
    }
    // also reject duplicates among aliased (removed) labels vs kept ones
    for(std::size_t k = 0; k < nalias; ++k)
        for(std::size_t b = 1; b < nb; ++b)
            if(alias_from[k] == t.block_label[b])
                lowering_error("duplicate label");

    auto block_of = [&](std::string_view n) -> std::uint16_t
    {
        n = resolve(n);
        if(n == "HALT") return halt16;
        for(std::size_t b = 0; b < nb; ++b)
            if(t.block_label[b] == n) return static_cast<std::uint16_t>(b);
        lowering_error("jump to an undefined label");
        return halt16;
    };

    t.nb           = nb;
    t.straightline = nb == 1;

    // ---- 3. instructions, fall-through jumps, successors ------------------
    for(std::size_t b = 0; b < nb; ++b)
    {
        const std::size_t s_end = b + 1 < nb ? bstart[b + 1] - 1 : last + 1;
        Block&            blk   = t.block[b];
        blk.i0                  = static_cast<std::uint16_t>(t.ni);
        std::size_t ncomp       = 0;
        std::size_t flops       = 0;
        bool        last_jump   = false;
        for(std::size_t i = bstart[b]; i < s_end; ++i)
        {
            const Stmt& s = raw.stmt[i];
            if(s.kind == SK::label) continue; // merged labels inside a run
            Instr in;
            in.kind = s.kind;
            in.op   = s.op;
            in.ety  = s.ety;
            in.ity  = s.ity;
            in.acc  = s.acc;
            in.cmp  = s.cmp;
            in.dst  = s.dst;
            in.a0   = s.a0;
            in.na   = s.na;
            in.c0   = s.c0;
            in.nc   = s.nc;
            if(s.kind == SK::jump || s.kind == SK::pjump)
            {
                in.t0 = static_cast<std::uint16_t>(t.nt);
                in.nt = s.nl;
                for(std::size_t k = 0; k < s.nl; ++k)
                    t.target[t.nt++] = block_of(raw.label[s.l0 + k]);
                last_jump = true;
            }
            else
            {
                ++ncomp;
                last_jump = false;
                if(ty_float(s.ety) && s.op != Op::mov && s.op != Op::cmov &&
                   s.op != Op::sct && s.op != Op::slt && s.op != Op::assert_fp)
                    flops += (s.op == Op::fma || s.op == Op::fms) ? 2 : 1;
            }
            t.instr[t.ni++] = in;
        }
        if(!t.straightline && !last_jump)
        {
            Instr in;
            in.kind          = SK::jump;
            in.t0            = static_cast<std::uint16_t>(t.nt);
            in.nt            = 1;
            t.target[t.nt++] = b + 1 < nb ? static_cast<std::uint16_t>(b + 1) :
                                            halt16;
            t.instr[t.ni++]  = in;
        }
        blk.ni  = static_cast<std::uint16_t>(t.ni - blk.i0);
        blk.ops   = static_cast<std::uint16_t>(ncomp + 1);
        blk.flops = static_cast<std::uint16_t>(flops);
        // successors: targets of the last instruction (straight-line: none)
        blk.s0 = static_cast<std::uint16_t>(t.nsucc);
        if(!t.straightline)
        {
            const Instr& term = t.instr[t.ni - 1];
            for(std::size_t k = 0; k < term.nt; ++k)
                t.succ[t.nsucc++] = t.target[term.t0 + k];
        }
        blk.ns = static_cast<std::uint16_t>(t.nsucc - blk.s0);
    }
    if(nb > 64) lowering_error("more than 64 basic blocks");

    // ---- 4. registers ---------------------------------------------------------
    for(std::size_t o = 0; o < raw.no; ++o) t.opnd[o] = raw.opnd[o];
    t.no = raw.no;
    for(std::size_t c = 0; c < raw.nc; ++c) t.konst[c] = raw.konst[c];
    t.nc = raw.nc;
    auto slot_of = [&](Operand& op) -> std::uint16_t
    {
        for(std::size_t r = 0; r < t.nr; ++r)
            if(t.reg[r].name == op.name)
            {
                Reg& g = t.reg[r];
                if(ty_size(g.ty) != ty_size(op.ty))
                    lowering_error("register used with types of different sizes");
                if(g.ty != op.ty) g.ntypes = 2;
                return static_cast<std::uint16_t>(r);
            }
        if(t.nr >= max_regs) lowering_error("more than 256 registers");
        Reg& g   = t.reg[t.nr];
        g.name   = op.name;
        g.ty     = op.ty;
        g.width  = static_cast<std::uint8_t>(ty_size(op.ty));
        g.ntypes = 1;
        return static_cast<std::uint16_t>(t.nr++);
    };
    // inputs first, then outputs, then operands in statement order
    for(std::size_t o = in0; o < raw.no; ++o)
    {
        if(t.opnd[o].lit) lowering_error("kernel input/output named like a literal");
        t.opnd[o].slot = slot_of(t.opnd[o]);
    }
    for(std::size_t o = 0; o < in0; ++o)
        if(!t.opnd[o].lit) t.opnd[o].slot = slot_of(t.opnd[o]);
    for(std::size_t k = 0; k < NIn; ++k)
    {
        t.in_slot[k] = t.opnd[in0 + k].slot;
        t.in_ty[k]   = t.opnd[in0 + k].ty;
        t.reg[t.in_slot[k]].input = true;
    }
    for(std::size_t k = 0; k < NOut; ++k)
    {
        t.out_slot[k] = t.opnd[out0 + k].slot;
        t.out_ty[k]   = t.opnd[out0 + k].ty;
        t.reg[t.out_slot[k]].output = true;
        t.outputs.set(t.out_slot[k]);
    }
    t.nin  = NIn;
    t.nout = NOut;
    for(std::size_t i = 0; i < t.ni; ++i)
        if(t.instr[i].kind == SK::compute)
        {
            const Operand& d = t.opnd[t.instr[i].dst];
            if(d.lit) lowering_error("destination named like a literal");
            t.reg[d.slot].written = true;
        }
    for(std::size_t i = 0; i < t.ni; ++i)
        for(std::size_t k = 0; k < t.instr[i].na; ++k)
        {
            const Operand& a = t.opnd[t.instr[i].a0 + k];
            if(!a.lit && !t.reg[a.slot].written && !t.reg[a.slot].input)
                lowering_error("register read but never written");
        }
    for(std::size_t r = 0; r < t.nr; ++r) t.state_size += t.reg[r].width;
    if(!t.straightline) t.state_size += t.index_bytes; // the PC register

    // ParmJump variables have the PC type (core.hpp:782-785)
    for(std::size_t i = 0; i < t.ni; ++i)
        if(t.instr[i].kind == SK::pjump)
        {
            const Ty vt = t.opnd[t.instr[i].a0].ty;
            if(ty_size(vt) != t.index_bytes || ty_float(vt) || ty_signed(vt))
                lowering_error("ParmJump variable must have the index type");
        }

    // ---- 5. liveness ---------------------------------------------------------
    for(std::size_t b = 0; b < nb; ++b)
    {
        Block& blk = t.block[b];
        for(std::size_t i = blk.i0; i < blk.i0 + blk.ni; ++i)
        {
            const Instr& in = t.instr[i];
            for(std::size_t k = 0; k < in.na; ++k)
            {
                const Operand& a = t.opnd[in.a0 + k];
                if(!a.lit && !blk.def.test(a.slot)) blk.use.set(a.slot);
            }
            if(in.kind == SK::compute) blk.def.set(t.opnd[in.dst].slot);
        }
    }
    for(bool changed = true; changed;)
    {
        changed = false;
        for(std::size_t bb = nb; bb-- > 0;)
        {
            Block& blk = t.block[bb];
            Bits   out{};
            if(blk.ns == 0) out = t.outputs;
            for(std::size_t k = 0; k < blk.ns; ++k)
            {
                const std::uint16_t s = t.succ[blk.s0 + k];
                out = out | (s == halt16 ? t.outputs : t.block[s].live_in);
            }
            const Bits in = blk.use | minus(out, blk.def);
            if(!(in == blk.live_in))
            {
                blk.live_in = in;
                changed     = true;
            }
        }
    }

    // ---- 6. reconvergence order and loops ----------------------------------
    {
        std::array<std::uint8_t, 64> state{}; // 0 new, 1 on stack, 2 done
        std::array<std::uint8_t, 64> post{};
        std::size_t                  npost = 0;
        std::array<std::uint8_t, 64> stack{}, next{};
        std::array<std::uint64_t, 64> back{}; // back[h]: sources of back edges to h
        std::size_t sp = 0;
        stack[sp++]    = 0;
        state[0]       = 1;
        while(sp)
        {
            const std::uint8_t b   = stack[sp - 1];
            const Block&       blk = t.block[b];
            if(next[b] < blk.ns)
            {
                const std::uint16_t s = t.succ[blk.s0 + next[b]++];
                if(s == halt16) continue;
                if(state[s] == 0)
                {
                    state[s]    = 1;
                    stack[sp++] = static_cast<std::uint8_t>(s);
                }
                else if(state[s] == 1)
                    back[s] |= std::uint64_t{1} << b;
            }
            else
            {
                state[b]      = 2;
                post[npost++] = b;
                --sp;
            }
        }
        std::array<std::uint8_t, 64> rpo{};
        std::size_t                  nr = 0;
        for(std::size_t k = npost; k-- > 0;) rpo[nr++] = post[k];
        for(std::size_t b = 0; b < nb; ++b)
            if(state[b] == 0) rpo[nr++] = static_cast<std::uint8_t>(b); // unreachable
        // natural loop bodies: blocks reaching a back-edge source without
        // passing through the header
        for(std::size_t h = 0; h < nb; ++h)
        {
            if(!back[h]) continue;
            std::uint64_t body = 0, work = back[h] & ~(std::uint64_t{1} << h);
            while(work)
            {
                std::size_t b = 0;
                while(!((work >> b) & 1u)) ++b;
                work &= ~(std::uint64_t{1} << b);
                if((body >> b) & 1u) continue;
                body |= std::uint64_t{1} << b;
                for(std::size_t p = 0; p < nb; ++p) // predecessors of b
                {
                    const Block& pb = t.block[p];
                    for(std::size_t k = 0; k < pb.ns; ++k)
                        if(t.succ[pb.s0 + k] == b && p != h && !((body >> p) & 1u))
                            work |= std::uint64_t{1} << p;
                }
            }
            t.head[t.nhead]   = static_cast<std::uint8_t>(h);
            t.body[t.nhead++] = body;
        }
        // scheduling order: blocks outside every loop first (each runs at
        // most once per thread, so running them early frees lanes at loop
        // exits), then loop blocks, both in reverse post-order
        std::uint64_t in_loop = 0;
        for(std::size_t k = 0; k < t.nhead; ++k)
            in_loop |= t.body[k] | (std::uint64_t{1} << t.head[k]);
        std::size_t r = 0;
        for(int pass = 0; pass < 2; ++pass)
            for(std::size_t k = 0; k < nr; ++k)
                if((((in_loop >> rpo[k]) & 1u) != 0) == (pass == 1)) t.order[r++] = rpo[k];
        for(std::size_t k = 0; k < nb; ++k) t.rank[t.order[k]] = static_cast<std::uint8_t>(k);
    }

    // ---- 7. structural hash -----------------------------------------------------
    std::uint64_t h = 0xCBF29CE484222325ull;
    h               = fnv(h, t.kname);
    for(std::size_t r = 0; r < t.nr; ++r)
    {
        h = fnv(h, t.reg[r].name);
        h = fnv(h, static_cast<std::uint64_t>(t.reg[r].ty));
    }
    for(std::size_t i = 0; i < t.ni; ++i)
    {
        const Instr& in = t.instr[i];
        h = fnv(h, static_cast<std::uint64_t>(in.kind) |
                       (static_cast<std::uint64_t>(in.op) << 8) |
                       (static_cast<std::uint64_t>(in.ety) << 16) |
                       (static_cast<std::uint64_t>(in.ity) << 24));
        for(std::size_t k = 0; k < in.na; ++k)
        {
            const Operand& a = t.opnd[in.a0 + k];
            h = fnv(h, a.lit ? a.bits : 0x8000000000000000ull | a.slot);
        }
        for(std::size_t k = 0; k < in.nt; ++k) h = fnv(h, t.target[in.t0 + k]);
    }
    for(std::size_t k = 0; k < t.nin; ++k) h = fnv(h, t.in_slot[k]);
    for(std::size_t k = 0; k < t.nout; ++k) h = fnv(h, t.out_slot[k]);
    t.hash = h;
    return t;
}

/// @brief The lowered table of kernel @p K.
template <Kernel K>
inline constexpr auto table_of = lower<K>();
} // namespace loom::vm::ir
