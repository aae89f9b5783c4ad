// -*- C++ -*-
/// @file bse.hpp
/// @brief ORACLE (test infrastructure only): a scalar, one-thread-at-a-time
/// interpreter of kernel_blank IR — the restatement of the reference scalar
/// backend (src/vm/bse.hpp:139-319) used to check the gfx950 backend.
///
/// It shares nothing with the product lowering except the IR types that
/// spell a kernel.  The statement list is walked at run time into a linear
/// program; control follows the reference block semantics directly:
///  - a block is the run of statements between labels (core.hpp:977-979);
///  - every statement of a block executes; the block ends at the next label
///    and transfers to the target of the block's last statement if that is a
///    jump, else falls through (ensure_last_jumps, core.hpp:804-829);
///  - labels before the first instruction are START, labels after the last
///    are HALT (trim, core.hpp:849-867);
///  - Jump sets the target (core.hpp:756-769); ParmJump indexes its label
///    list by the variable (core.hpp:771-791): two labels select on
///    nonzero, more labels halt on an out-of-range index;
///  - Compute stores Op(args) into Dst (core.hpp:741-753) with the op
///    contract of ops.hpp.
#pragma once

#include "ops.hpp"

#include <vm/core.hpp>
#include <vm/operators.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace loom_oracle
{
enum class Code : int
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
    copysign
};

template <typename T>
Ty
type_code()
{
    if constexpr(std::is_same_v<T, float>) return Ty::f32;
    else if constexpr(std::is_same_v<T, double>) return Ty::f64;
    else if constexpr(std::is_signed_v<T>)
        return sizeof(T) == 1 ? Ty::i8 : sizeof(T) == 2 ? Ty::i16 : sizeof(T) == 4 ? Ty::i32 : Ty::i64;
    else
        return sizeof(T) == 1 ? Ty::u8 : sizeof(T) == 2 ? Ty::u16 : sizeof(T) == 4 ? Ty::u32 : Ty::u64;
}

struct Arg
{
    bool          literal = false;
    int           reg     = -1;
    Ty            ty      = Ty::u64;
    std::uint64_t bits    = 0;
    std::string   name;
};

struct Stmt
{
    enum Kind
    {
        compute,
        label,
        jump,
        pjump
    } kind = compute;
    Code                       code{};
    Ty                         ety = Ty::u64; // evaluation type
    Ty                         ity = Ty::u64; // index / predicate type
    Code                       acc{}, cmp{};  // pacc
    Arg                        dst;
    std::vector<Arg>           args;
    std::vector<std::string>   labels; // label name (label) or targets (jumps)
    std::vector<std::uint64_t> konst;  // sct constants (bits)
    std::vector<int>           target; // resolved jump targets (-1 = HALT, -2 = START)
};

// ---- mapping of operation types to codes -------------------------------------
template <typename O>
struct op_of;
#define LOOM_ORACLE_OP(FN, C)                    \
    template <typename T>                        \
    struct op_of<FN<T>>                          \
    {                                            \
        static constexpr Code code = Code::C;    \
        using type                 = T;          \
    };
LOOM_ORACLE_OP(std::plus, add)
LOOM_ORACLE_OP(std::minus, sub)
LOOM_ORACLE_OP(std::multiplies, mul)
LOOM_ORACLE_OP(std::divides, div)
LOOM_ORACLE_OP(std::modulus, mod)
LOOM_ORACLE_OP(std::negate, neg)
LOOM_ORACLE_OP(std::equal_to, eq)
LOOM_ORACLE_OP(std::not_equal_to, ne)
LOOM_ORACLE_OP(std::less, lt)
LOOM_ORACLE_OP(std::greater, gt)
LOOM_ORACLE_OP(std::less_equal, le)
LOOM_ORACLE_OP(std::greater_equal, ge)
LOOM_ORACLE_OP(std::logical_and, land)
LOOM_ORACLE_OP(std::logical_or, lor)
LOOM_ORACLE_OP(std::logical_not, lnot)
LOOM_ORACLE_OP(std::bit_and, band)
LOOM_ORACLE_OP(std::bit_or, bor)
LOOM_ORACLE_OP(std::bit_xor, bxor)
LOOM_ORACLE_OP(std::bit_not, bnot)
LOOM_ORACLE_OP(loom::vm::ops::lshift, shl)
LOOM_ORACLE_OP(loom::vm::ops::rshift, shr)
LOOM_ORACLE_OP(loom::vm::ops::abs, abs)
LOOM_ORACLE_OP(loom::vm::ops::rsqrt, rsqrt)
LOOM_ORACLE_OP(loom::vm::ops::min, min)
LOOM_ORACLE_OP(loom::vm::ops::max, max)
LOOM_ORACLE_OP(loom::vm::ops::round, round)
LOOM_ORACLE_OP(loom::vm::ops::ceil, ceil)
LOOM_ORACLE_OP(loom::vm::ops::floor, floor)
LOOM_ORACLE_OP(loom::vm::ops::trunc, trunc)
LOOM_ORACLE_OP(loom::vm::ops::sin, sin)
LOOM_ORACLE_OP(loom::vm::ops::cos, cos)
LOOM_ORACLE_OP(loom::vm::ops::sqrt, sqrt)
LOOM_ORACLE_OP(loom::vm::ops::copysign, copysign)
#undef LOOM_ORACLE_OP
template <>
struct op_of<std::identity>
{
    static constexpr Code code = Code::mov;
    using type                 = void;
};
template <>
struct op_of<loom::vm::ops::cmov>
{
    static constexpr Code code = Code::cmov;
    using type                 = void;
};
template <>
struct op_of<loom::vm::ops::fma>
{
    static constexpr Code code = Code::fma;
    using type                 = void;
};
template <>
struct op_of<loom::vm::ops::fms>
{
    static constexpr Code code = Code::fms;
    using type                 = void;
};
template <>
struct op_of<loom::vm::ops::assert_fp>
{
    static constexpr Code code = Code::assert_fp;
    using type                 = void;
};
template <typename T, typename I, typename L>
struct op_of<loom::vm::ops::sct<T, I, L>>
{
    static constexpr Code code = Code::sct;
    using type                 = T;
    using index                = I;
};
template <typename T, typename I>
struct op_of<loom::vm::ops::slt<T, I>>
{
    static constexpr Code code = Code::slt;
    using type                 = T;
    using index                = I;
};
template <typename T, typename I, template <typename> typename A, template <typename> typename C>
struct op_of<loom::vm::ops::pacc<T, I, A, C>>
{
    static constexpr Code code = Code::pacc;
    using type                 = T;
    using index                = I;
    static constexpr Code acc  = op_of<A<I>>::code;
    static constexpr Code cmp  = op_of<C<T>>::code;
};

// ---- statement collection ------------------------------------------------------
template <typename V>
Arg
make_arg()
{
    using T = typename V::type;
    Arg a;
    a.name = std::string(V::name::view);
    a.ty   = type_code<T>();
    if(auto v = lit::parse<T>(V::name::view))
    {
        a.literal = true;
        T tv      = *v;
        std::memcpy(&a.bits, &tv, sizeof(T));
    }
    return a;
}

template <typename S>
struct walk
{
    static void
    run(std::vector<Stmt>&)
    {
    } // macro_inv and other markers
};

template <typename... Ts>
struct walk<loom::type_sequence<Ts...>>
{
    static void
    run(std::vector<Stmt>& out)
    {
        (walk<Ts>::run(out), ...);
    }
};

template <typename N>
struct walk<loom::vm::Label<N>>
{
    static void
    run(std::vector<Stmt>& out)
    {
        Stmt s;
        s.kind = Stmt::label;
        s.labels.emplace_back(N::view);
        out.push_back(std::move(s));
    }
};

template <typename N>
struct walk<loom::vm::Jump<loom::vm::Label<N>>>
{
    static void
    run(std::vector<Stmt>& out)
    {
        Stmt s;
        s.kind = Stmt::jump;
        s.labels.emplace_back(N::view);
        out.push_back(std::move(s));
    }
};

template <typename V, typename... Ns>
struct walk<loom::vm::ParmJump<V, loom::type_sequence<loom::vm::Label<Ns>...>>>
{
    static void
    run(std::vector<Stmt>& out)
    {
        Stmt s;
        s.kind = Stmt::pjump;
        s.args.push_back(make_arg<V>());
        (s.labels.emplace_back(Ns::view), ...);
        out.push_back(std::move(s));
    }
};

template <typename T, typename... Cs>
void
sct_constants(Stmt& s, loom::type_sequence<Cs...>)
{
    (
        [&]
        {
            auto v = lit::parse<T>(Cs::view);
            if(!v) throw std::runtime_error("sct constant is not a number");
            T             tv = *v;
            std::uint64_t b  = 0;
            std::memcpy(&b, &tv, sizeof(T));
            s.konst.push_back(b);
        }(),
        ...);
}

template <typename O, typename D, typename... As>
struct walk<loom::vm::Compute<O, D, As...>>
{
    static void
    run(std::vector<Stmt>& out)
    {
        using tr = op_of<O>;
        Stmt s;
        s.kind = Stmt::compute;
        s.code = tr::code;
        s.dst  = make_arg<D>();
        // a destination whose name parses as a number is a literal, which
        // the reference rejects at compile time (core.hpp:966-970)
        if(s.dst.literal)
            throw std::runtime_error("destination named like a literal: " + s.dst.name);
        (s.args.push_back(make_arg<As>()), ...);
        if constexpr(!std::is_void_v<typename tr::type>)
            s.ety = type_code<typename tr::type>();
        else if constexpr(tr::code == Code::cmov)
            s.ety = s.args.at(1).ty;
        else
            s.ety = s.args.at(0).ty;
        if constexpr(tr::code == Code::sct || tr::code == Code::slt || tr::code == Code::pacc)
            s.ity = type_code<typename tr::index>();
        if constexpr(tr::code == Code::pacc)
        {
            s.acc = tr::acc;
            s.cmp = tr::cmp;
        }
        if constexpr(tr::code == Code::sct)
            sct_constants<typename tr::type>(s, typename O::list{});
        out.push_back(std::move(s));
    }
};

// ---- program -------------------------------------------------------------------
struct Program
{
    std::string              name;
    std::vector<Stmt>        code;
    std::vector<std::string> reg_name;
    std::vector<int>         in_reg, out_reg;
    std::vector<Ty>          in_ty, out_ty;
    std::vector<int>         start_zero; // every register (zeroed per thread)
    std::size_t              first = 0;  // first non-label statement
};

template <typename... Vs>
void
io_regs(loom::type_sequence<Vs...>, std::vector<Arg>& out)
{
    (out.push_back(make_arg<Vs>()), ...);
}

template <typename K>
Program
build_program()
{
    Program p;
    p.name = std::string(K::nameview);
    walk<typename K::statements>::run(p.code);

    std::map<std::string, int> reg;
    auto intern = [&](Arg& a)
    {
        if(a.literal) return;
        auto it = reg.find(a.name);
        if(it == reg.end())
        {
            it = reg.emplace(a.name, static_cast<int>(p.reg_name.size())).first;
            p.reg_name.push_back(a.name);
        }
        a.reg = it->second;
    };
    std::vector<Arg> ins, outs;
    io_regs(typename K::input_vars{}, ins);
    io_regs(typename K::output_vars{}, outs);
    for(auto& a: ins)
    {
        intern(a);
        p.in_reg.push_back(a.reg);
        p.in_ty.push_back(a.ty);
    }
    for(auto& a: outs)
    {
        intern(a);
        p.out_reg.push_back(a.reg);
        p.out_ty.push_back(a.ty);
    }
    for(auto& s: p.code)
    {
        if(s.kind == Stmt::compute) intern(s.dst);
        for(auto& a: s.args) intern(a);
    }

    // label positions; leading labels are START, trailing labels are HALT
    std::size_t first = p.code.size(), last = 0;
    for(std::size_t i = 0; i < p.code.size(); ++i)
        if(p.code[i].kind != Stmt::label)
        {
            if(first == p.code.size()) first = i;
            last = i;
        }
    p.first = first;
    std::map<std::string, int> pos;
    for(std::size_t i = 0; i < p.code.size(); ++i)
        if(p.code[i].kind == Stmt::label)
        {
            const std::string& n = p.code[i].labels[0];
            const int where      = i < first ? -2 : (i > last ? -1 : static_cast<int>(i));
            if(!pos.emplace(n, where).second) throw std::runtime_error("duplicate label " + n);
        }
    for(auto& s: p.code)
        if(s.kind == Stmt::jump || s.kind == Stmt::pjump)
            for(auto& l: s.labels)
            {
                if(l == "HALT")
                    s.target.push_back(-1);
                else if(l == "START")
                    s.target.push_back(-2);
                else
                {
                    auto it = pos.find(l);
                    if(it == pos.end()) throw std::runtime_error("undefined label " + l);
                    s.target.push_back(it->second);
                }
            }
    return p;
}

// ---- evaluation ------------------------------------------------------------------
template <typename V>
V
reg_as(std::uint64_t bits)
{
    V v;
    std::memcpy(&v, &bits, sizeof(V));
    return v;
}

template <typename V>
std::uint64_t
as_bits(V v)
{
    std::uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(V));
    return b;
}

/// Value of @p a converted to E.
template <typename E>
E
value(const Arg& a, const std::uint64_t* R)
{
    const std::uint64_t b = a.literal ? a.bits : R[a.reg];
    switch(a.ty)
    {
    case Ty::u8: return convert<E>(reg_as<std::uint8_t>(b));
    case Ty::u16: return convert<E>(reg_as<std::uint16_t>(b));
    case Ty::u32: return convert<E>(reg_as<std::uint32_t>(b));
    case Ty::u64: return convert<E>(reg_as<std::uint64_t>(b));
    case Ty::i8: return convert<E>(reg_as<std::int8_t>(b));
    case Ty::i16: return convert<E>(reg_as<std::int16_t>(b));
    case Ty::i32: return convert<E>(reg_as<std::int32_t>(b));
    case Ty::i64: return convert<E>(reg_as<std::int64_t>(b));
    case Ty::f32: return convert<E>(reg_as<float>(b));
    default: return convert<E>(reg_as<double>(b));
    }
}

/// Store @p v (type E) into the register of @p d converted to d's type.
template <typename E>
void
store(const Arg& d, E v, std::uint64_t* R)
{
    std::uint64_t& slot = R[d.reg];
    switch(d.ty)
    {
    case Ty::u8: slot = as_bits(convert<std::uint8_t>(v)); break;
    case Ty::u16: slot = as_bits(convert<std::uint16_t>(v)); break;
    case Ty::u32: slot = as_bits(convert<std::uint32_t>(v)); break;
    case Ty::u64: slot = as_bits(convert<std::uint64_t>(v)); break;
    case Ty::i8: slot = as_bits(convert<std::int8_t>(v)); break;
    case Ty::i16: slot = as_bits(convert<std::int16_t>(v)); break;
    case Ty::i32: slot = as_bits(convert<std::int32_t>(v)); break;
    case Ty::i64: slot = as_bits(convert<std::int64_t>(v)); break;
    case Ty::f32: slot = as_bits(convert<float>(v)); break;
    default: slot = as_bits(convert<double>(v)); break;
    }
}

template <typename E>
bool
compare(Code c, E a, E b)
{
    switch(c)
    {
    case Code::eq: return a == b;
    case Code::ne: return a != b;
    case Code::lt: return a < b;
    case Code::gt: return a > b;
    case Code::le: return a <= b;
    default: return a >= b;
    }
}

/// Integer arithmetic modulo 2^bits, with the promoted width of C++ for
/// shifts (32 bits for types up to 4 bytes).
template <typename E>
void
eval_int(const Stmt& s, std::uint64_t* R)
{
    using U       = std::make_unsigned_t<E>;
    using W       = std::conditional_t<(sizeof(E) <= 4), std::uint32_t, std::uint64_t>;
    const E a     = s.args.empty() ? E(0) : value<E>(s.args[0], R);
    const E b     = s.args.size() > 1 ? value<E>(s.args[1], R) : E(0);
    const U ua    = static_cast<U>(a), ub = static_cast<U>(b);
    const int wb  = sizeof(E) <= 4 ? 32 : 64;
    switch(s.code)
    {
    case Code::add: store<E>(s.dst, static_cast<E>(static_cast<U>(ua + ub)), R); break;
    case Code::sub: store<E>(s.dst, static_cast<E>(static_cast<U>(ua - ub)), R); break;
    case Code::mul: store<E>(s.dst, static_cast<E>(static_cast<U>(W(ua) * W(ub))), R); break;
    case Code::div:
        if(b == 0)
            store<E>(s.dst, static_cast<E>(~U(0)), R);
        else if(std::is_signed_v<E> && b == E(-1))
            store<E>(s.dst, static_cast<E>(static_cast<U>(U(0) - ua)), R);
        else
            store<E>(s.dst, static_cast<E>(a / b), R);
        break;
    case Code::mod:
        if(b == 0)
            store<E>(s.dst, a, R);
        else if(std::is_signed_v<E> && b == E(-1))
            store<E>(s.dst, E(0), R);
        else
            store<E>(s.dst, static_cast<E>(a % b), R);
        break;
    case Code::neg: store<E>(s.dst, static_cast<E>(static_cast<U>(U(0) - ua)), R); break;
    case Code::band: store<E>(s.dst, static_cast<E>(ua & ub), R); break;
    case Code::bor: store<E>(s.dst, static_cast<E>(ua | ub), R); break;
    case Code::bxor: store<E>(s.dst, static_cast<E>(ua ^ ub), R); break;
    case Code::bnot: store<E>(s.dst, static_cast<E>(static_cast<U>(~ua)), R); break;
    case Code::shl:
        store<E>(s.dst, static_cast<E>(static_cast<U>(W(ua) << (W(ub) % wb))), R);
        break;
    case Code::shr:
        if constexpr(std::is_signed_v<E>)
        {
            using SW = std::make_signed_t<W>;
            store<E>(s.dst, static_cast<E>(SW(a) >> (W(ub) % wb)), R);
        }
        else
            store<E>(s.dst, static_cast<E>(W(ua) >> (W(ub) % wb)), R);
        break;
    case Code::abs:
        store<E>(s.dst, (std::is_signed_v<E> && a < 0) ? static_cast<E>(static_cast<U>(U(0) - ua)) : a, R);
        break;
    case Code::min: store<E>(s.dst, b < a ? b : a, R); break;
    case Code::max: store<E>(s.dst, a < b ? b : a, R); break;
    default: throw std::runtime_error("integer operation not in the contract");
    }
}

template <typename E>
void
eval_float(const Stmt& s, std::uint64_t* R)
{
    const E a = value<E>(s.args[0], R);
    const E b = s.args.size() > 1 ? value<E>(s.args[1], R) : E(0);
    switch(s.code)
    {
    case Code::add: store<E>(s.dst, a + b, R); break;
    case Code::sub: store<E>(s.dst, a - b, R); break;
    case Code::mul: store<E>(s.dst, a * b, R); break;
    case Code::div: store<E>(s.dst, a / b, R); break;
    case Code::neg: store<E>(s.dst, -a, R); break;
    case Code::abs: store<E>(s.dst, std::fabs(a), R); break;
    case Code::rsqrt: store<E>(s.dst, E(1) / std::sqrt(a), R); break;
    case Code::sqrt: store<E>(s.dst, std::sqrt(a), R); break;
    case Code::min: store<E>(s.dst, b < a ? b : a, R); break;
    case Code::max: store<E>(s.dst, a < b ? b : a, R); break;
    case Code::round: store<E>(s.dst, std::round(a), R); break;
    case Code::ceil: store<E>(s.dst, std::ceil(a), R); break;
    case Code::floor: store<E>(s.dst, std::floor(a), R); break;
    case Code::trunc: store<E>(s.dst, std::trunc(a), R); break;
    case Code::copysign: store<E>(s.dst, std::copysign(a, b), R); break;
    case Code::sin: store<E>(s.dst, sin_poly(a), R); break;
    case Code::cos: store<E>(s.dst, cos_poly(a), R); break;
    default: throw std::runtime_error("floating-point operation not in the contract");
    }
}

/// Everything not specific to integers or floats, in evaluation type E.
template <typename E>
void
eval_typed(const Stmt& s, std::uint64_t* R, std::uint32_t& err)
{
    switch(s.code)
    {
    case Code::eq:
    case Code::ne:
    case Code::lt:
    case Code::gt:
    case Code::le:
    case Code::ge:
        store<bool>(s.dst, compare<E>(s.code, value<E>(s.args[0], R), value<E>(s.args[1], R)), R);
        return;
    case Code::land:
        store<bool>(s.dst, value<E>(s.args[0], R) != E(0) && value<E>(s.args[1], R) != E(0), R);
        return;
    case Code::lor:
        store<bool>(s.dst, value<E>(s.args[0], R) != E(0) || value<E>(s.args[1], R) != E(0), R);
        return;
    case Code::lnot: store<bool>(s.dst, !(value<E>(s.args[0], R) != E(0)), R); return;
    case Code::cmov:
    {
        // the condition is tested in its own type
        const bool c = value<double>(s.args[0], R) != 0.0 ||
                       (s.args[0].ty != Ty::f32 && s.args[0].ty != Ty::f64 &&
                        value<std::uint64_t>(s.args[0], R) != 0);
        store<E>(s.dst, c ? value<E>(s.args[1], R) : value<E>(s.args[2], R), R);
        return;
    }
    case Code::fma:
    case Code::fms:
        if constexpr(std::is_floating_point_v<E>)
        {
            const E z = value<E>(s.args[2], R);
            store<E>(s.dst,
                     std::fma(value<E>(s.args[0], R), value<E>(s.args[1], R),
                              s.code == Code::fma ? z : -z),
                     R);
            return;
        }
        else
            throw std::runtime_error("fma needs floating point");
    case Code::sct:
    {
        std::uint64_t i = 0;
        switch(s.ity)
        {
        case Ty::u32: i = value<std::uint32_t>(s.args[0], R); break;
        case Ty::u64: i = value<std::uint64_t>(s.args[0], R); break;
        case Ty::u16: i = value<std::uint16_t>(s.args[0], R); break;
        case Ty::u8: i = value<std::uint8_t>(s.args[0], R); break;
        default: i = static_cast<std::uint64_t>(value<std::int64_t>(s.args[0], R)); break;
        }
        store<E>(s.dst, i < s.konst.size() ? reg_as<E>(s.konst[i]) : E(0), R);
        return;
    }
    case Code::slt:
    {
        std::uint64_t i = 0;
        switch(s.ity)
        {
        case Ty::u32: i = value<std::uint32_t>(s.args[0], R); break;
        case Ty::u64: i = value<std::uint64_t>(s.args[0], R); break;
        case Ty::u16: i = value<std::uint16_t>(s.args[0], R); break;
        case Ty::u8: i = value<std::uint8_t>(s.args[0], R); break;
        default: i = static_cast<std::uint64_t>(value<std::int64_t>(s.args[0], R)); break;
        }
        store<E>(s.dst, i + 1 < s.args.size() ? value<E>(s.args[i + 1], R) : E(0), R);
        return;
    }
    case Code::pacc:
    {
        const bool p = value<std::uint64_t>(s.args[0], R) != 0;
        const bool c = compare<E>(s.cmp, value<E>(s.args[1], R), value<E>(s.args[2], R));
        store<bool>(s.dst, s.acc == Code::land ? (p && c) : (p || c), R);
        return;
    }
    case Code::assert_fp:
    {
        const E v = value<E>(s.args[0], R);
        if constexpr(std::is_floating_point_v<E>)
            if(!std::isfinite(v)) err |= 1u;
        store<E>(s.dst, v, R);
        return;
    }
    default:
        if constexpr(std::is_floating_point_v<E>)
            eval_float<E>(s, R);
        else
            eval_int<E>(s, R);
    }
}

inline void
eval_compute(const Stmt& s, std::uint64_t* R, std::uint32_t& err)
{
    if(s.code == Code::mov)
    {
        // the source keeps its own type; converted on store
        const Arg& a = s.args[0];
        switch(a.ty)
        {
        case Ty::u8: store(s.dst, value<std::uint8_t>(a, R), R); break;
        case Ty::u16: store(s.dst, value<std::uint16_t>(a, R), R); break;
        case Ty::u32: store(s.dst, value<std::uint32_t>(a, R), R); break;
        case Ty::u64: store(s.dst, value<std::uint64_t>(a, R), R); break;
        case Ty::i8: store(s.dst, value<std::int8_t>(a, R), R); break;
        case Ty::i16: store(s.dst, value<std::int16_t>(a, R), R); break;
        case Ty::i32: store(s.dst, value<std::int32_t>(a, R), R); break;
        case Ty::i64: store(s.dst, value<std::int64_t>(a, R), R); break;
        case Ty::f32: store(s.dst, value<float>(a, R), R); break;
        default: store(s.dst, value<double>(a, R), R); break;
        }
        return;
    }
    switch(s.ety)
    {
    case Ty::u8: eval_typed<std::uint8_t>(s, R, err); break;
    case Ty::u16: eval_typed<std::uint16_t>(s, R, err); break;
    case Ty::u32: eval_typed<std::uint32_t>(s, R, err); break;
    case Ty::u64: eval_typed<std::uint64_t>(s, R, err); break;
    case Ty::i8: eval_typed<std::int8_t>(s, R, err); break;
    case Ty::i16: eval_typed<std::int16_t>(s, R, err); break;
    case Ty::i32: eval_typed<std::int32_t>(s, R, err); break;
    case Ty::i64: eval_typed<std::int64_t>(s, R, err); break;
    case Ty::f32: eval_typed<float>(s, R, err); break;
    default: eval_typed<double>(s, R, err); break;
    }
}

/// Run one thread to HALT, calling on_block(first) after every block with the
/// index of the block's first statement; on_block returns false to stop.
/// A block ends where control reaches a label or the end of the code (a
/// pending jump is taken there, else control falls through).  R holds one
/// 64-bit word per register (low bytes significant), zeroed by the caller
/// except for the inputs.  Returns the number of executed statements
/// (labels excluded).
template <typename OnBlock>
std::uint64_t
run_blocks(const Program& p, std::uint64_t* R, std::uint32_t& err, std::uint64_t max_statements,
           OnBlock&& on_block)
{
    const std::size_t n       = p.code.size();
    std::size_t       ip      = p.first;
    int               pending = -3; // -3: none
    long              first   = -1; // first statement of the current block
    std::uint64_t     count   = 0;
    auto              go      = [&](int t) -> bool
    {
        if(t == -1) return false;
        ip = t == -2 ? p.first : static_cast<std::size_t>(t);
        return true;
    };
    auto end_block = [&]() -> bool
    {
        if(first < 0) return true;
        const long f = first;
        first        = -1;
        return on_block(static_cast<std::size_t>(f));
    };
    for(;;)
    {
        if(ip >= n)
        {
            if(!end_block()) return count;
            if(pending != -3 && go(pending))
            {
                pending = -3;
                continue;
            }
            return count;
        }
        const Stmt& s = p.code[ip];
        if(s.kind == Stmt::label)
        {
            if(!end_block()) return count;
            if(pending != -3)
            {
                const int t = pending;
                pending     = -3;
                if(!go(t)) return count;
                continue;
            }
            ++ip;
            continue;
        }
        if(++count > max_statements)
        {
            err |= 2u;
            return count;
        }
        if(first < 0) first = static_cast<long>(ip);
        if(s.kind == Stmt::compute)
        {
            eval_compute(s, R, err);
            pending = -3; // a later statement overrides an earlier jump
        }
        else if(s.kind == Stmt::jump)
            pending = s.target[0];
        else
        {
            const Arg&          v  = s.args[0];
            const std::uint64_t iv = v.literal ? v.bits : R[v.reg];
            std::uint64_t       x  = iv;
            if(ty_bytes(v.ty) < 8) x &= (std::uint64_t(1) << (8 * ty_bytes(v.ty))) - 1;
            if(s.target.size() == 2)
                pending = x != 0 ? s.target[1] : s.target[0];
            else
                pending = x < s.target.size() ? s.target[x] : -1;
        }
        ++ip;
    }
}

/// Run one thread to HALT (run_blocks without a per-block hook).
inline std::uint64_t
run_thread(const Program& p, std::uint64_t* R, std::uint32_t& err,
           std::uint64_t max_statements = ~std::uint64_t(0))
{
    return run_blocks(p, R, err, max_statements, [](std::size_t) { return true; });
}

/// Names of the labels directly before statement @p i (the labels that name
/// the block starting there), "START" first when @p i is the first statement.
inline std::string
labels_before(const Program& p, std::size_t i)
{
    std::string out = i == p.first ? "START" : "";
    for(std::size_t j = i; j > 0 && p.code[j - 1].kind == Stmt::label; --j)
        for(const auto& l: p.code[j - 1].labels) out += (out.empty() ? "" : " ") + l;
    return out;
}
} // namespace loom_oracle
