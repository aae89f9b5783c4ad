// -*- C++ -*-
/// @file literal.hpp
/// @brief Constant-evaluated parser for IR literals.
///
/// A variable whose name parses as a number of the variable's type is a
/// literal (reference core.hpp:443-456).  The value is defined by the
/// reference's own parsers, which are deliberately inexact for floating point
/// (common/string.hpp:312-352): integer part + fraction digits times
/// 10^-len, then times 10^exp, all in the target type, with powers computed
/// by squaring (ext/include/static_math/detail/cmath.inl:48-56, 618-628).
/// Integers accept "0x"/"0b" prefixes (string.hpp:286-306).  This file
/// restates those rules so literal values match bit for bit.
#pragma once

#include <bit>
#include <cstdint>
#include <optional>
#include <string_view>
#include <type_traits>

namespace loom::vm::literal
{
/// @brief `c2i` (string.hpp:188-203): digit value of @p c in @p base.
constexpr int
digit(char c, unsigned base) noexcept
{
    if(base == 16)
    {
        if(c >= 'a' && c <= 'f') return c - 87;
        if(c >= 'A' && c <= 'F') return c - 55;
    }
    if(c >= '0' && c <= '9')
    {
        const int v = c - 48;
        if(static_cast<unsigned>(v) < base) return v;
    }
    return -1;
}

/// @brief `parse_num` (string.hpp:222-238) in wrapping 64-bit arithmetic:
/// digits are accumulated from the right as digit·base^k.
constexpr std::optional<std::uint64_t>
parse_digits(std::string_view s, unsigned base) noexcept
{
    if(s.empty()) return std::nullopt;
    std::uint64_t ret = 0, power = 1;
    for(std::size_t k = s.size(); k-- > 0;)
    {
        const int d = digit(s[k], base);
        if(d < 0) return std::nullopt;
        ret += static_cast<std::uint64_t>(d) * power;
        power *= base;
    }
    return ret;
}

/// @brief static_math `pow_helper`: exponentiation by squaring, evaluated in
/// the floating type @p T (cmath.inl:48-56).
template <typename T>
constexpr T
pow_sq(T x, long e) noexcept
{
    if(e == 0) return static_cast<T>(1);
    if(e % 2 == 0) return pow_sq<T>(x * x, e / 2);
    return x * pow_sq<T>(x * x, (e - 1) / 2);
}

/// @brief static_math `pow` for an integer exponent (cmath.inl:618-628).
template <typename T>
constexpr T
pow_int(T x, long e) noexcept
{
    if(e == 0) return static_cast<T>(1);
    if(e > 0) return pow_sq<T>(x, e);
    return static_cast<T>(1) / pow_sq<T>(x, -e);
}

/// @brief `s2f_fn` (string.hpp:312-352).
///
/// A literal reads "[sign] whole [. fraction] [e|E [sign] exponent]".  Its
/// value is built in T as the reference builds it: whole as an integer
/// converted to T, plus fraction-as-integer times 10^-(fraction digits),
/// times 10^exponent, times the sign.  Where the reference finds the first
/// '.' and the first 'e'/'E' of the whole text, this parser cuts the text
/// at the exponent marker first and looks for '.' only before it; the
/// accepted set is the same, because a '.' after the marker leaves the
/// marker among the whole-part digits there and among the exponent digits
/// here, and both parsers then reject the text.  Empty pieces are skipped:
/// "e5" is 0 and "1e" is 1, as in the reference.
template <typename T>
constexpr std::optional<T>
parse_float(std::string_view text) noexcept
{
    T sign = static_cast<T>(1);
    if(!text.empty() && (text.front() == '-' || text.front() == '+'))
    {
        if(text.front() == '-') sign = static_cast<T>(-1);
        text.remove_prefix(1);
    }
    if(text.empty()) return std::nullopt;

    std::size_t marker = 0;
    while(marker < text.size() && text[marker] != 'e' && text[marker] != 'E') ++marker;
    std::string_view       mantissa = text.substr(0, marker);
    const std::string_view exponent =
        marker < text.size() ? text.substr(marker + 1) : std::string_view{};

    std::string_view fraction{};
    if(const std::size_t dot = mantissa.find('.'); dot != std::string_view::npos)
    {
        fraction = mantissa.substr(dot + 1);
        mantissa = mantissa.substr(0, dot);
    }

    T value{};
    if(!mantissa.empty())
    {
        const auto digits = parse_digits(mantissa, 10);
        if(!digits) return std::nullopt;
        value = static_cast<T>(static_cast<long>(*digits));
    }
    if(!fraction.empty())
    {
        const auto digits = parse_digits(fraction, 10);
        if(!digits) return std::nullopt;
        const long places = static_cast<long>(static_cast<int>(fraction.size()));
        value += static_cast<T>(static_cast<long>(*digits)) *
                 pow_int<T>(static_cast<T>(10), -places);
    }
    if(!exponent.empty())
    {
        std::string_view mag  = exponent;
        long             sign_e = 1;
        if(mag.front() == '-' || mag.front() == '+')
        {
            if(mag.front() == '-') sign_e = -1;
            mag.remove_prefix(1);
        }
        if(mag.empty()) return std::nullopt;
        const auto digits = parse_digits(mag, 10);
        if(!digits) return std::nullopt;
        value *= pow_int<T>(static_cast<T>(10), static_cast<long>(*digits) * sign_e);
    }
    return value * sign;
}

/// @brief `s2i_base_prefix_fn` (string.hpp:286-306) for an integer of type
/// @p Int, returned as the value's bit pattern in 64 bits.
template <typename Int>
constexpr std::optional<Int>
parse_int(std::string_view s) noexcept
{
    if(s.empty()) return std::nullopt;
    bool neg = false;
    if constexpr(std::is_signed_v<Int>)
    {
        if(s[0] == '-' || s[0] == '+')
        {
            neg = s[0] == '-';
            s   = s.substr(1);
        }
    }
    unsigned base = 10;
    if(s.substr(0, 2) == "0x" || s.substr(0, 2) == "0X")
    {
        base = 16;
        s    = s.substr(2);
    }
    else if(s.substr(0, 2) == "0b" || s.substr(0, 2) == "0B")
    {
        base = 2;
        s    = s.substr(2);
    }
    auto v = parse_digits(s, base);
    if(!v) return std::nullopt;
    std::uint64_t u = *v;
    if(neg) u = ~u + 1;
    using U = std::make_unsigned_t<Int>;
    return static_cast<Int>(static_cast<U>(u));
}

/// @brief `s2n<T>` (string.hpp:437-447, 470-471).
template <typename T>
constexpr std::optional<T>
parse(std::string_view s) noexcept
{
    if constexpr(std::is_floating_point_v<T>)
        return parse_float<T>(s);
    else
        return parse_int<T>(s);
}

/// @brief Bit pattern of a literal, zero-extended to 64 bits.
template <typename T>
constexpr std::uint64_t
bits_of(T v) noexcept
{
    if constexpr(sizeof(T) == 8)
        return std::bit_cast<std::uint64_t>(v);
    else if constexpr(sizeof(T) == 4)
        return std::bit_cast<std::uint32_t>(v);
    else if constexpr(sizeof(T) == 2)
        return std::bit_cast<std::uint16_t>(v);
    else
        return std::bit_cast<std::uint8_t>(v);
}

/// The reference's own checks of its float parser (src/tests/string.cpp:40-56):
/// these texts parse to the value the compiler gives the same literal, and
/// three texts are rejected.  Every translation unit that includes this
/// header re-checks them at compile time.
namespace checks
{
static_assert(parse_float<float>("0").value() == 0.0f);
static_assert(parse_float<float>("1").value() == 1.0f);
static_assert(parse_float<float>("1234.567").value() == 1234.567f);
static_assert(parse_float<float>("1e2").value() == 1e2f);
static_assert(parse_float<float>("0.5e2").value() == 0.5e2f);
static_assert(parse_float<float>("0.5e-2").value() == 0.5e-2f);
static_assert(parse_float<float>("1e-8").value() == 1e-8f);
static_assert(parse_float<double>("0").value() == 0.0);
static_assert(parse_float<double>("-1234.567").value() == -1234.567);
static_assert(parse_float<double>("-1e2").value() == -1e2);
static_assert(parse_float<double>("-0.5e2").value() == -0.5e2);
static_assert(parse_float<double>("-0.5e-2").value() == -0.5e-2);
static_assert(parse_float<double>("-2430.5e4").value() == -2430.5e4);
static_assert(!parse_float<float>("e334g").has_value());
static_assert(!parse_float<float>("-").has_value());
static_assert(!parse_float<float>("+").has_value());
} // namespace checks
} // namespace loom::vm::literal
