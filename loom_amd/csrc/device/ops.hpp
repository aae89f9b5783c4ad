// -*- C++ -*-
/// @file ops.hpp
/// @brief gfx950 implementation of the op contract.
///
/// Every operation of the IR has one definition shared with the CPU oracle
/// (which restates it independently).  Where the reference leaves behaviour
/// undefined (shift counts, integer division by zero, float→int overflow,
/// out-of-range table indices) the contract fixes a value.  All floating
/// point is IEEE round-to-nearest with no contraction (the library is built
/// with -ffp-contract=off); fma/fms are fused; sin/cos are the polynomial of
/// the reference SIMD backends (simd.hpp:955-1100, degree 15/16 for float,
/// 33/32 for double, Horner form with fused steps, simd.hpp:194-203).
#pragma once

#include <vm/lower.hpp>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#ifndef LOOM_DEVICE
#define LOOM_DEVICE __device__ __forceinline__

/// A compile-time value of the lowered tables used as an immediate: forces
/// constant evaluation so device code never loads table entries from memory.
#define LOOM_CX(...) (::std::integral_constant<decltype(__VA_ARGS__), (__VA_ARGS__)>::value)
#endif

namespace loom::hipvm
{
namespace ir = loom::vm::ir;
using ir::Op;
using ir::Ty;

template <Ty t>
struct ctype_of;
template <>
struct ctype_of<Ty::u8>
{
    using type = std::uint8_t;
};
template <>
struct ctype_of<Ty::u16>
{
    using type = std::uint16_t;
};
template <>
struct ctype_of<Ty::u32>
{
    using type = std::uint32_t;
};
template <>
struct ctype_of<Ty::u64>
{
    using type = std::uint64_t;
};
template <>
struct ctype_of<Ty::i8>
{
    using type = std::int8_t;
};
template <>
struct ctype_of<Ty::i16>
{
    using type = std::int16_t;
};
template <>
struct ctype_of<Ty::i32>
{
    using type = std::int32_t;
};
template <>
struct ctype_of<Ty::i64>
{
    using type = std::int64_t;
};
template <>
struct ctype_of<Ty::f32>
{
    using type = float;
};
template <>
struct ctype_of<Ty::f64>
{
    using type = double;
};

template <Ty t>
using ctype = typename ctype_of<t>::type;

/// @brief Numeric conversion of the contract: integer↔integer is modular,
/// integer→float rounds to nearest, float→integer truncates toward zero and
/// saturates to the target range with NaN → 0.
template <typename To, typename From>
LOOM_DEVICE To
conv(From v)
{
    if constexpr(std::is_same_v<To, From>)
        return v;
    else if constexpr(std::is_floating_point_v<From> &&
                      std::is_integral_v<To>)
    {
        constexpr int  bits = int(sizeof(To) * 8);
        constexpr bool sgn  = std::is_signed_v<To>;
        // lo = min(To), hi = max(To) + 1; both are powers of two (or 0) and
        // therefore exact in From
        constexpr From lo = sgn ? -From(std::uint64_t(1) << (bits - 1)) : From(0);
        constexpr From hi = sgn ? From(std::uint64_t(1) << (bits - 1)) :
                            (bits == 64 ? From(18446744073709551616.0) :
                                          From(std::uint64_t(1) << (bits % 64)));
        using U = std::make_unsigned_t<To>;
        constexpr U umax = U(~U(0));
        constexpr To vmax = sgn ? To(umax >> 1) : To(umax);
        constexpr To vmin = sgn ? To(U(umax >> 1) + U(1)) : To(0);
        if(v != v) return To(0);
        if(v <= lo) return vmin;
        if(v >= hi) return vmax;
        return To(v);
    }
    else if constexpr(std::is_integral_v<From> && std::is_integral_v<To>)
        return To(std::make_unsigned_t<To>(v));
    else
        return To(v);
}

template <typename T>
LOOM_DEVICE T
fma_(T x, T y, T z)
{
    if constexpr(std::is_same_v<T, float>)
        return __builtin_fmaf(x, y, z);
    else
        return __builtin_fma(x, y, z);
}

/// @brief Correctly rounded f32 square root, the hardware estimate refined by
/// its two neighbours: v_sqrt_f32 is within one ulp, and of s - 1 ulp, s and
/// s + 1 ulp the residuals x - s'·s (exact with fma) pick the nearest root.
/// The compiler's correctly rounded sqrt does the same on x scaled by 2^32
/// when x < 2^-96 and returns x itself for ±0, +inf and NaN; for x = ±0,
/// +inf and every x >= 2^-96 the refinement alone gives the same bits, so
/// sqrt_rn takes it when no active lane of the wave holds another value
/// (a wave-uniform branch, no per-lane divergence), else the compiler's.
LOOM_DEVICE float
sqrt_refined(float x)
{
    const float         s  = __builtin_amdgcn_sqrtf(x);
    const std::uint32_t si = __builtin_bit_cast(std::uint32_t, s);
    const float         sm = __builtin_bit_cast(float, si - 1u);
    const float         sp = __builtin_bit_cast(float, si + 1u);
    const float         rm = __builtin_fmaf(-sm, s, x);
    const float         rp = __builtin_fmaf(-sp, s, x);
    float               r  = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : r;
}

LOOM_DEVICE float
sqrt_rn(float x)
{
    const bool other = !(x >= 0x1p-96f || x == 0.0f); // NaN, negative, tiny
    if(__builtin_amdgcn_ballot_w64(other) == 0) return sqrt_refined(x);
    return __builtin_sqrtf(x);
}

/// @brief Correctly rounded 1 / (correctly rounded sqrt(x)) in f32.  The
/// compiler's division is the Newton sequence y = rcp(d), y += y(1 - dy),
/// q = n·y, q += y(n - dq) twice, wrapped in v_div_scale (a no-op unless an
/// operand or an intermediate leaves the normal range), v_div_fmas (an fma
/// when nothing was scaled) and v_div_fixup (the special operands).  For
/// x in [2^-96, 2^126] the root d lies in [2^-48, 2^63], none of that applies
/// and the bare sequence gives the same bits; the wave takes it when every
/// active lane's x is in that range.
LOOM_DEVICE float
rsqrt_rn(float x)
{
    const bool other = !(x >= 0x1p-96f && x <= 0x1p126f); // also NaN
    if(__builtin_amdgcn_ballot_w64(other) == 0)
    {
        const float d  = sqrt_refined(x);
        float       y  = __builtin_amdgcn_rcpf(d);
        y              = __builtin_fmaf(__builtin_fmaf(-d, y, 1.0f), y, y);
        float       q  = y;
        q              = __builtin_fmaf(__builtin_fmaf(-d, q, 1.0f), y, q);
        return __builtin_fmaf(__builtin_fmaf(-d, q, 1.0f), y, q);
    }
    return 1.0f / __builtin_sqrtf(x);
}

/// @brief Polynomial sine of the reference SIMD backends, valid on [-π, π].
template <typename T>
LOOM_DEVICE T
poly_sin(T x)
{
    const T x2 = x * x;
    T       p;
    if constexpr(std::is_same_v<T, float>)
    {
        p = fma_(x2, -7.37066278281678179293546235379257048e-13f,
                 1.6047844633018114428263064708844387e-10f);
        p = fma_(p, x2, -2.5051870883490902518590017165648945e-8f);
        p = fma_(p, x2, 2.75573164212929639596445204076151988e-6f);
        p = fma_(p, x2, -0.000198412698232225093689107237380802637f);
        p = fma_(p, x2, 0.00833333333327592139676057538106745659f);
        p = fma_(p, x2, -0.166666666666659653164780128859839316f);
        p = fma_(p, x2, 0.99999999999999985783856947280351013f);
    }
    else
    {
        p = fma_(x2, 1.1309232145353729608313042496143056e-37,
                 -1.21592124202981493418420109807105932e-34);
        p = fma_(p, x2, 1.13099507260187617461053670403005966e-31);
        p = fma_(p, x2, -9.18368981471916803573077288615983239e-29);
        p = fma_(p, x2, 6.4469502842428826022931795893278841e-26);
        p = fma_(p, x2, -3.86817017063038299631930987571173923e-23);
        p = fma_(p, x2, 1.95729410633912564384267263771198416e-20);
        p = fma_(p, x2, -8.22063524662432971120933376991172232e-18);
        p = fma_(p, x2, 2.81145725434552076319376582531091405e-15);
        p = fma_(p, x2, -7.64716373181981647590109722915908442e-13);
        p = fma_(p, x2, 1.60590438368216145993923769996397832e-10);
        p = fma_(p, x2, -2.50521083854417187750521083848472947e-8);
        p = fma_(p, x2, 2.75573192239858906525573192239844806e-6);
        p = fma_(p, x2, -0.000198412698412698412698412698412698391);
        p = fma_(p, x2, 0.00833333333333333333333333333333333333);
        p = fma_(p, x2, -0.166666666666666666666666666666666667);
        p = fma_(p, x2, 1.);
    }
    return x * p;
}

/// @brief Polynomial cosine of the reference SIMD backends, valid on [-π, π].
template <typename T>
LOOM_DEVICE T
poly_cos(T x)
{
    const T x2 = x * x;
    T       p;
    if constexpr(std::is_same_v<T, float>)
    {
        p = fma_(x2, 4.60900737685258733987906059894416831e-14f,
                 -1.14629048996344469555493302470243688e-11f);
        p = fma_(p, x2, 2.08765619601383980877940138085106396e-9f);
        p = fma_(p, x2, -2.75573163935355075808996593111486352e-7f);
        p = fma_(p, x2, 0.0000248015872774439536215582612189964674f);
        p = fma_(p, x2, -0.00138888888887731722447063574018337351f);
        p = fma_(p, x2, 0.041666666666663887964497659102665977f);
        p = fma_(p, x2, -0.49999999999999974308667070690271678f);
        p = fma_(p, x2, 0.99999999999999999608981951072301546f);
    }
    else
    {
        p = fma_(x2, 3.72999385718624695747058735680768987e-36,
                 -3.76931659747655711635436902424567497e-33);
        p = fma_(p, x2, 3.27988537078655827882448844587725832e-30);
        p = fma_(p, x2, -2.47959624821373207705239218354539419e-27);
        p = fma_(p, x2, 1.61173757105463581590277799117611453e-24);
        p = fma_(p, x2, -8.89679139244973290214501480503565676e-22);
        p = fma_(p, x2, 4.11031762331216359173758562466598535e-19);
        p = fma_(p, x2, -1.5619206968586226447944923010120297e-16);
        p = fma_(p, x2, 4.77947733238738529742623766654306629e-14);
        p = fma_(p, x2, -1.14707455977297247138516241570198547e-11);
        p = fma_(p, x2, 2.08767569878680989792100899936785891e-9);
        p = fma_(p, x2, -2.75573192239858906525573192229758788e-7);
        p = fma_(p, x2, 0.0000248015873015873015873015873015852479);
        p = fma_(p, x2, -0.00138888888888888888888888888888888864);
        p = fma_(p, x2, 0.0416666666666666666666666666666666667);
        p = fma_(p, x2, -0.5);
        p = fma_(p, x2, 1.);
    }
    return p;
}

/// @brief Shift count mask: counts are taken modulo the promoted width
/// (32 bits for types of at most 4 bytes, 64 bits otherwise).
template <typename T>
inline constexpr unsigned shift_mask = sizeof(T) <= 4 ? 31u : 63u;

template <typename T>
using promoted_u = std::conditional_t<sizeof(T) <= 4, std::uint32_t, std::uint64_t>;

/// @brief Binary/unary arithmetic of the contract in the evaluation type @p T.
template <Op op, typename T>
LOOM_DEVICE auto
arith(T a, T b = T{})
{
    if constexpr(op == Op::add || op == Op::sub || op == Op::mul)
    {
        if constexpr(std::is_integral_v<T>)
        {
            using U = std::make_unsigned_t<T>;
            const U x = U(a), y = U(b);
            if constexpr(op == Op::add) return T(U(x + y));
            if constexpr(op == Op::sub) return T(U(x - y));
            if constexpr(op == Op::mul) return T(U(promoted_u<T>(x) * promoted_u<T>(y)));
        }
        else
        {
            if constexpr(op == Op::add) return T(a + b);
            if constexpr(op == Op::sub) return T(a - b);
            if constexpr(op == Op::mul) return T(a * b);
        }
    }
    else if constexpr(op == Op::div)
    {
        if constexpr(std::is_integral_v<T>)
        {
            if(b == T(0)) return T(~T(0));
            if constexpr(std::is_signed_v<T>)
                if(b == T(-1)) return T(std::make_unsigned_t<T>(0) - std::make_unsigned_t<T>(a));
            return T(a / b);
        }
        else
            return T(a / b);
    }
    else if constexpr(op == Op::mod)
    {
        static_assert(std::is_integral_v<T>, "modulus needs an integer type");
        if(b == T(0)) return a;
        if constexpr(std::is_signed_v<T>)
            if(b == T(-1)) return T(0);
        return T(a % b);
    }
    else if constexpr(op == Op::neg)
    {
        if constexpr(std::is_integral_v<T>)
            return T(std::make_unsigned_t<T>(0) - std::make_unsigned_t<T>(a));
        else
            return T(-a);
    }
    else if constexpr(op == Op::band || op == Op::bor || op == Op::bxor ||
                      op == Op::bnot)
    {
        static_assert(std::is_integral_v<T>, "bit operation needs an integer type");
        if constexpr(op == Op::band) return T(a & b);
        if constexpr(op == Op::bor) return T(a | b);
        if constexpr(op == Op::bxor) return T(a ^ b);
        if constexpr(op == Op::bnot) return T(~a);
    }
    else if constexpr(op == Op::shl || op == Op::shr)
    {
        static_assert(std::is_integral_v<T>, "shift needs an integer type");
        const unsigned s = unsigned(std::make_unsigned_t<T>(b)) & shift_mask<T>;
        if constexpr(op == Op::shl)
            return T(promoted_u<T>(std::make_unsigned_t<T>(a)) << s);
        else if constexpr(std::is_signed_v<T>)
        {
            using S = std::conditional_t<sizeof(T) <= 4, std::int32_t, std::int64_t>;
            return T(S(a) >> s);
        }
        else
            return T(promoted_u<T>(a) >> s);
    }
    else if constexpr(op == Op::eq)
        return a == b;
    else if constexpr(op == Op::ne)
        return a != b;
    else if constexpr(op == Op::lt)
        return a < b;
    else if constexpr(op == Op::gt)
        return a > b;
    else if constexpr(op == Op::le)
        return a <= b;
    else if constexpr(op == Op::ge)
        return a >= b;
    else if constexpr(op == Op::land)
        return (a != T(0)) && (b != T(0));
    else if constexpr(op == Op::lor)
        return (a != T(0)) || (b != T(0));
    else if constexpr(op == Op::lnot)
        return !(a != T(0));
    else if constexpr(op == Op::abs)
    {
        if constexpr(std::is_same_v<T, float>)
            return __builtin_fabsf(a);
        else if constexpr(std::is_same_v<T, double>)
            return __builtin_fabs(a);
        else if constexpr(std::is_signed_v<T>)
            return a < T(0) ? T(std::make_unsigned_t<T>(0) - std::make_unsigned_t<T>(a)) : a;
        else
            return a;
    }
    else if constexpr(op == Op::min)
        return (b < a) ? b : a;
    else if constexpr(op == Op::max)
        return (a < b) ? b : a;
    else
    {
        static_assert(std::is_floating_point_v<T>, "operation needs a floating-point type");
        if constexpr(op == Op::rsqrt)
        {
            if constexpr(std::is_same_v<T, float>)
                return rsqrt_rn(a);
            else
                return 1.0 / __builtin_sqrt(a);
        }
        else if constexpr(op == Op::sqrt)
        {
            if constexpr(std::is_same_v<T, float>)
                return sqrt_rn(a);
            else
                return __builtin_sqrt(a);
        }
        else if constexpr(op == Op::round)
        {
            if constexpr(std::is_same_v<T, float>)
                return __builtin_roundf(a);
            else
                return __builtin_round(a);
        }
        else if constexpr(op == Op::ceil)
        {
            if constexpr(std::is_same_v<T, float>)
                return __builtin_ceilf(a);
            else
                return __builtin_ceil(a);
        }
        else if constexpr(op == Op::floor)
        {
            if constexpr(std::is_same_v<T, float>)
                return __builtin_floorf(a);
            else
                return __builtin_floor(a);
        }
        else if constexpr(op == Op::trunc)
        {
            if constexpr(std::is_same_v<T, float>)
                return __builtin_truncf(a);
            else
                return __builtin_trunc(a);
        }
        else if constexpr(op == Op::copysign)
        {
            if constexpr(std::is_same_v<T, float>)
                return __builtin_copysignf(a, b);
            else
                return __builtin_copysign(a, b);
        }
        else if constexpr(op == Op::sin)
            return poly_sin<T>(a);
        else if constexpr(op == Op::cos)
            return poly_cos<T>(a);
        else
            static_assert(sizeof(T) == 0, "unhandled operation");
    }
}
} // namespace loom::hipvm
