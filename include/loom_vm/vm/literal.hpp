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
template <typename T>
constexpr std::optional<T>
parse_float(std::string_view word) noexcept
{
    T sign = static_cast<T>(1);
    if(!word.empty() && (word[0] == '-' || word[0] == '+'))
    {
        if(word[0] == '-') sign = static_cast<T>(-1);
        word = word.substr(1);
    }
    const std::string_view s = word;
    if(s.empty()) return std::nullopt;

    const std::size_t sep_pos = s.find('.');
    const std::size_t ep1     = s.find('e');
    const std::size_t ep2     = s.find('E');
    const std::size_t exp_pos = ep1 < ep2 ? ep1 : ep2;
    const bool        has_sep = sep_pos != std::string_view::npos;
    const bool        has_exp = exp_pos != std::string_view::npos;

    const std::string_view int_part =
        has_sep ? s.substr(0, sep_pos) : (has_exp ? s.substr(0, exp_pos) : s);
    const std::string_view frac_part =
        has_sep ? (has_exp ? s.substr(sep_pos + 1, exp_pos - sep_pos - 1) :
                             s.substr(sep_pos + 1)) :
                  std::string_view{};
    const std::string_view exp_part =
        has_exp ? s.substr(exp_pos + 1) : std::string_view{};

    T ret{};
    if(!int_part.empty())
    {
        const auto oi = parse_digits(int_part, 10);
        if(!oi) return std::nullopt;
        ret = static_cast<T>(static_cast<long>(*oi));
    }
    if(!frac_part.empty())
    {
        const auto oi = parse_digits(frac_part, 10);
        if(!oi) return std::nullopt;
        const T frac = static_cast<T>(static_cast<long>(*oi)) *
                       pow_int<T>(static_cast<T>(10),
                                  -static_cast<long>(static_cast<int>(
                                      frac_part.length())));
        ret += frac;
    }
    if(!exp_part.empty())
    {
        std::string_view e   = exp_part;
        long             esg = 1;
        if(e[0] == '-' || e[0] == '+')
        {
            if(e[0] == '-') esg = -1;
            e = e.substr(1);
        }
        if(e.empty()) return std::nullopt;
        const auto oi = parse_digits(e, 10);
        if(!oi) return std::nullopt;
        ret *= pow_int<T>(static_cast<T>(10), static_cast<long>(*oi) * esg);
    }
    return ret * sign;
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
} // namespace loom::vm::literal
