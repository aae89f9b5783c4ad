// -*- C++ -*-
/// @file ops.hpp
/// @brief ORACLE (test infrastructure only): scalar restatement of the IR op
/// contract and of the reference literal parsers.
///
/// Only tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() use
/// this code, as a checker.  It is written independently of the gfx950
/// lowering (loom_amd/csrc/device/ops.hpp) so that the two can be compared.
///
/// Sources restated:
///  - scalar op semantics: reference src/vm/operators.hpp:81-330 and the std
///    functors named in operators.hpp:20-50;
///  - sin/cos: the SIMD polynomial, src/vm/simd.hpp:955-1100 with the
///    Horner/fma nesting of simd.hpp:194-203;
///  - literal values: src/common/string.hpp:188-352 (c2i, parse_num,
///    parse_signed, detect_base, s2i_base_prefix_fn, s2f_fn) and
///    ext/include/static_math/detail/cmath.inl:48-56, 618-628 (pow).
/// Behaviour the reference leaves undefined is fixed by the op contract
/// (shift counts modulo the promoted width, integer x/0 = all ones, x%0 = x,
/// saturating float->integer conversion with NaN -> 0, out-of-range table
/// index -> 0, multiway jump index out of range -> HALT).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <optional>
#include <string_view>
#include <type_traits>

namespace loom_oracle
{
enum class Ty : int
{
    u8 = 0,
    u16,
    u32,
    u64,
    i8,
    i16,
    i32,
    i64,
    f32,
    f64
};

inline int
ty_bytes(Ty t)
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
    default: return 8;
    }
}

// ---- conversions -------------------------------------------------------------
template <typename To, typename From>
To
convert(From v)
{
    if constexpr(std::is_same_v<To, From>)
        return v;
    else if constexpr(std::is_floating_point_v<From> && std::is_integral_v<To>)
    {
        if(std::isnan(v)) return 0;
        // exact range test in long double is not needed: compare against the
        // powers of two bounding To, which are exact in From
        const long double x  = v;
        const long double lo = static_cast<long double>(std::numeric_limits<To>::min());
        const long double hi = std::is_signed_v<To> ?
                                   -lo :
                                   std::ldexp(1.0L, static_cast<int>(sizeof(To) * 8));
        if(x <= lo) return std::numeric_limits<To>::min();
        if(x >= hi) return std::numeric_limits<To>::max();
        return static_cast<To>(v);
    }
    else if constexpr(std::is_integral_v<From> && std::is_integral_v<To>)
    {
        using U = std::make_unsigned_t<To>;
        return static_cast<To>(static_cast<U>(v));
    }
    else
        return static_cast<To>(v);
}

// ---- polynomials (coefficient data of the reference, simd.hpp:968-1090) -----
inline float
sin_poly(float x)
{
    static const float k[8] = {0.99999999999999985783856947280351013f,
                               -0.166666666666659653164780128859839316f,
                               0.00833333333327592139676057538106745659f,
                               -0.000198412698232225093689107237380802637f,
                               2.75573164212929639596445204076151988e-6f,
                               -2.5051870883490902518590017165648945e-8f,
                               1.6047844633018114428263064708844387e-10f,
                               -7.37066278281678179293546235379257048e-13f};
    const float x2  = x * x;
    float       acc = std::fma(x2, k[7], k[6]);
    for(int i = 5; i >= 0; --i) acc = std::fma(acc, x2, k[i]);
    return x * acc;
}

inline float
cos_poly(float x)
{
    static const float k[9] = {0.99999999999999999608981951072301546f,
                               -0.49999999999999974308667070690271678f,
                               0.041666666666663887964497659102665977f,
                               -0.00138888888887731722447063574018337351f,
                               0.0000248015872774439536215582612189964674f,
                               -2.75573163935355075808996593111486352e-7f,
                               2.08765619601383980877940138085106396e-9f,
                               -1.14629048996344469555493302470243688e-11f,
                               4.60900737685258733987906059894416831e-14f};
    const float x2  = x * x;
    float       acc = std::fma(x2, k[8], k[7]);
    for(int i = 6; i >= 0; --i) acc = std::fma(acc, x2, k[i]);
    return acc;
}

inline double
sin_poly(double x)
{
    static const double k[17] = {1.,
                                 -0.166666666666666666666666666666666667,
                                 0.00833333333333333333333333333333333333,
                                 -0.000198412698412698412698412698412698391,
                                 2.75573192239858906525573192239844806e-6,
                                 -2.50521083854417187750521083848472947e-8,
                                 1.60590438368216145993923769996397832e-10,
                                 -7.64716373181981647590109722915908442e-13,
                                 2.81145725434552076319376582531091405e-15,
                                 -8.22063524662432971120933376991172232e-18,
                                 1.95729410633912564384267263771198416e-20,
                                 -3.86817017063038299631930987571173923e-23,
                                 6.4469502842428826022931795893278841e-26,
                                 -9.18368981471916803573077288615983239e-29,
                                 1.13099507260187617461053670403005966e-31,
                                 -1.21592124202981493418420109807105932e-34,
                                 1.1309232145353729608313042496143056e-37};
    const double x2  = x * x;
    double       acc = std::fma(x2, k[16], k[15]);
    for(int i = 14; i >= 0; --i) acc = std::fma(acc, x2, k[i]);
    return x * acc;
}

inline double
cos_poly(double x)
{
    static const double k[17] = {1.,
                                 -0.5,
                                 0.0416666666666666666666666666666666667,
                                 -0.00138888888888888888888888888888888864,
                                 0.0000248015873015873015873015873015852479,
                                 -2.75573192239858906525573192229758788e-7,
                                 2.08767569878680989792100899936785891e-9,
                                 -1.14707455977297247138516241570198547e-11,
                                 4.77947733238738529742623766654306629e-14,
                                 -1.5619206968586226447944923010120297e-16,
                                 4.11031762331216359173758562466598535e-19,
                                 -8.89679139244973290214501480503565676e-22,
                                 1.61173757105463581590277799117611453e-24,
                                 -2.47959624821373207705239218354539419e-27,
                                 3.27988537078655827882448844587725832e-30,
                                 -3.76931659747655711635436902424567497e-33,
                                 3.72999385718624695747058735680768987e-36};
    const double x2  = x * x;
    double       acc = std::fma(x2, k[16], k[15]);
    for(int i = 14; i >= 0; --i) acc = std::fma(acc, x2, k[i]);
    return acc;
}

// ---- literal parsing (string.hpp:188-352; cmath.inl:48-56, 618-628) ------------
namespace lit
{
inline int
digit_value(char c, unsigned base)
{
    if(base == 16 && c >= 'a' && c <= 'f') return c - 'a' + 10;
    if(base == 16 && c >= 'A' && c <= 'F') return c - 'A' + 10;
    if(c >= '0' && c <= '9' && static_cast<unsigned>(c - '0') < base) return c - '0';
    return -1;
}

/// digits accumulated right to left as d * base^k in 64-bit wrapping
/// arithmetic (parse_num, string.hpp:222-238)
inline std::optional<std::uint64_t>
digits(std::string_view s, unsigned base)
{
    if(s.empty()) return std::nullopt;
    std::uint64_t v = 0, w = 1;
    for(auto it = s.rbegin(); it != s.rend(); ++it)
    {
        const int d = digit_value(*it, base);
        if(d < 0) return std::nullopt;
        v += w * static_cast<std::uint64_t>(d);
        w *= base;
    }
    return v;
}

/// static_math pow by squaring in the type of x
template <typename F>
F
square_pow(F x, long e)
{
    if(e == 0) return F(1);
    if(e % 2 == 0) return square_pow<F>(x * x, e / 2);
    return x * square_pow<F>(x * x, (e - 1) / 2);
}

template <typename F>
F
ipow(F x, long e)
{
    return e == 0 ? F(1) : (e > 0 ? square_pow<F>(x, e) : F(1) / square_pow<F>(x, -e));
}

template <typename F>
std::optional<F>
parse_real(std::string_view w)
{
    F s = 1;
    if(!w.empty() && (w.front() == '+' || w.front() == '-'))
    {
        s = w.front() == '-' ? F(-1) : F(1);
        w.remove_prefix(1);
    }
    if(w.empty()) return std::nullopt;
    const auto dot = w.find('.');
    const auto e   = std::min(w.find('e'), w.find('E'));
    std::string_view ip, fp, ep;
    if(dot != std::string_view::npos)
    {
        ip = w.substr(0, dot);
        fp = e != std::string_view::npos ? w.substr(dot + 1, e - dot - 1) : w.substr(dot + 1);
    }
    else
        ip = e != std::string_view::npos ? w.substr(0, e) : w;
    if(e != std::string_view::npos) ep = w.substr(e + 1);

    F v = 0;
    if(!ip.empty())
    {
        auto d = digits(ip, 10);
        if(!d) return std::nullopt;
        v = static_cast<F>(static_cast<long>(*d));
    }
    if(!fp.empty())
    {
        auto d = digits(fp, 10);
        if(!d) return std::nullopt;
        v += static_cast<F>(static_cast<long>(*d)) * ipow<F>(F(10), -static_cast<long>(fp.size()));
    }
    if(!ep.empty())
    {
        long sg = 1;
        if(ep.front() == '+' || ep.front() == '-')
        {
            sg = ep.front() == '-' ? -1 : 1;
            ep.remove_prefix(1);
        }
        if(ep.empty()) return std::nullopt;
        auto d = digits(ep, 10);
        if(!d) return std::nullopt;
        v *= ipow<F>(F(10), sg * static_cast<long>(*d));
    }
    return v * s;
}

template <typename I>
std::optional<I>
parse_integer(std::string_view w)
{
    if(w.empty()) return std::nullopt;
    bool neg = false;
    if(std::is_signed_v<I> && (w.front() == '+' || w.front() == '-'))
    {
        neg = w.front() == '-';
        w.remove_prefix(1);
    }
    unsigned base = 10;
    if(w.size() >= 2 && w[0] == '0' && (w[1] == 'x' || w[1] == 'X'))
    {
        base = 16;
        w.remove_prefix(2);
    }
    else if(w.size() >= 2 && w[0] == '0' && (w[1] == 'b' || w[1] == 'B'))
    {
        base = 2;
        w.remove_prefix(2);
    }
    auto d = digits(w, base);
    if(!d) return std::nullopt;
    std::uint64_t u = neg ? (0 - *d) : *d;
    return static_cast<I>(static_cast<std::make_unsigned_t<I>>(u));
}

template <typename V>
std::optional<V>
parse(std::string_view w)
{
    if constexpr(std::is_floating_point_v<V>)
        return parse_real<V>(w);
    else
        return parse_integer<V>(w);
}
} // namespace lit
} // namespace loom_oracle
