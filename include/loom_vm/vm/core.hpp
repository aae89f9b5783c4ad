// -*- C++ -*-
/// @file core.hpp
/// @brief IR entities of the Loom SPMD VM, source-compatible with the
/// reference front-end (src/vm/core.hpp:36-133, 939-1031).
///
/// A kernel is a type `kernel_blank<Name, Inputs, Statements, Outputs>` built
/// from `Variable`, `Compute`, `Label`, `Jump` and `ParmJump` types by the
/// macros of `vm/macros.hpp`.  Unlike the reference, this header performs no
/// type-level analysis: the types are only data.  All analysis (label
/// optimisation, block splitting, register inference, successors, liveness)
/// happens in constant-evaluated functions over flat arrays in
/// `vm/lower.hpp`, which keeps compile time linear in the kernel size.
#pragma once

#include "lists.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <string_view>
#include <type_traits>
#include <utility>

namespace loom
{
/// @brief Compile-time character sequence: the type-level name of registers,
/// labels and kernels (reference: common/tmp.hpp `char_sequence`).
template <char... Cs>
struct char_sequence
{
    static constexpr char             data[sizeof...(Cs) + 1] = {Cs..., '\0'};
    static constexpr std::string_view view{data, sizeof...(Cs)};
};

template <typename T>
struct is_char_sequence : std::false_type
{
};

template <char... Cs>
struct is_char_sequence<char_sequence<Cs...>> : std::true_type
{
};

template <typename T>
concept CharSequence = is_char_sequence<T>::value;

/// @brief String literal usable as a class-type template parameter.
template <std::size_t N>
struct string_literal
{
    char chars[N]{};

    constexpr string_literal(const char (&s)[N]) noexcept
    {
        for(std::size_t i = 0; i < N; ++i) chars[i] = s[i];
    }
};

namespace detail
{
template <string_literal S, typename Idx>
struct literal_to_cs;

template <string_literal S, std::size_t... I>
struct literal_to_cs<S, std::index_sequence<I...>>
{
    using type = char_sequence<S.chars[I]...>;
};
} // namespace detail

/// @brief String literal → char_sequence (reference: `loom::sl2cs`).
template <string_literal S>
using sl2cs = typename detail::literal_to_cs<
    S,
    std::make_index_sequence<sizeof(S.chars) - 1>>::type;

/// @brief string_view of a char_sequence.
template <CharSequence S>
inline constexpr std::string_view cs2sv = S::view;

/// @brief Plain type list (reference: `loom::type_sequence`).
template <typename... Ts>
struct type_sequence
{
    static constexpr std::size_t size = sizeof...(Ts);
};

/// @brief A two-type list (reference: `loom::type_pair`).
template <typename A, typename B>
using type_pair = type_sequence<A, B>;

namespace detail
{
template <std::size_t S>
struct uint_of_size;
template <>
struct uint_of_size<1>
{
    using type = std::uint8_t;
};
template <>
struct uint_of_size<2>
{
    using type = std::uint16_t;
};
template <>
struct uint_of_size<4>
{
    using type = std::uint32_t;
};
template <>
struct uint_of_size<8>
{
    using type = std::uint64_t;
};
} // namespace detail

/// @brief Unsigned integer of @p S bytes (reference: common/tmp.hpp:2021).
template <std::size_t S>
using uint_of_size_t = typename detail::uint_of_size<S>::type;
} // namespace loom

namespace loom::vm
{
/// @brief A register (or literal) reference: a fundamental type and a name
/// (reference core.hpp:38-44).  A name that parses as a number of @p Type is a
/// literal (core.hpp:443-456).
template <typename Type, CharSequence Name>
struct Variable
{
    using type                                 = Type;
    using name                                 = Name;
    static constexpr std::string_view nameview = Name::view;
};

template <typename V>
using vartype = typename V::type;

template <typename V>
using varname = typename V::name;

/// @brief `Dst = Op(Args...)` (reference core.hpp:74-82).
template <typename Op, typename Dst, typename... Args>
struct Compute
{
    using var    = Dst;
    using action = Op;
    using args   = type_sequence<Args...>;
};

/// @brief Basic-block label (reference core.hpp:87-91).
template <CharSequence Name>
struct Label
{
    using name = Name;
};

/// @brief Jump target that terminates the thread (core.hpp:94).
using HaltLabel  = Label<sl2cs<"HALT">>;
/// @brief Jump target that restarts the thread (core.hpp:97).
using StartLabel = Label<sl2cs<"START">>;

/// @brief Multiway jump: `pc = Labels[Var]` (core.hpp:117-123).
template <typename Var, typename Labels>
struct ParmJump
{
    using var    = Var;
    using labels = Labels;
};

/// @brief Unconditional jump (core.hpp:128-133).
template <typename Target>
struct Jump
{
    using target = Target;
};

/// @brief Macro declaration marker, kept for pretty printing only
/// (core.hpp:905-916).  Not a statement.
template <CharSequence Name, typename... Parameters>
struct macro_decl
{
    static constexpr std::string_view nameview = Name::view;
    using parameters                           = type_sequence<Parameters...>;
};

/// @brief Macro invocation marker (core.hpp:918-924).  Not a statement.
template <CharSequence MacroName, typename... Args>
struct macro_inv
{
    static constexpr std::string_view nameview = MacroName::view;
    using args                                 = type_sequence<Args...>;
};

namespace detail
{
template <typename L>
struct head_of;
template <typename T, typename... Ts>
struct head_of<type_sequence<T, Ts...>>
{
    using type = T;
};
} // namespace detail

/// @brief Kernel definition (reference core.hpp:939-1031): name, input
/// variables, statements (possibly nested lists from inserted macros) and
/// output variables.  `index_t` (the program-counter type) is the unsigned
/// integer of the size of the first input (core.hpp:950).
template <CharSequence Name,
          typename InputVars,
          typename Statements,
          typename OutputVars>
    requires(InputVars::size > 0 && Statements::size > 0 &&
             OutputVars::size > 0)
struct kernel_blank
{
    using name        = Name;
    using input_vars  = InputVars;
    using statements  = Statements;
    using output_vars = OutputVars;
    using index_t =
        uint_of_size_t<sizeof(vartype<typename detail::head_of<InputVars>::type>)>;

    static constexpr std::string_view nameview = Name::view;
    static constexpr index_t          halt_pos = static_cast<index_t>(-1);
};

template <typename K>
struct is_kernel : std::false_type
{
};

template <typename... Ts>
struct is_kernel<kernel_blank<Ts...>> : std::true_type
{
};

/// @brief A kernel ready to lower (reference core.hpp:1046-1047).
template <typename T>
concept Kernel = is_kernel<T>::value;
} // namespace loom::vm
