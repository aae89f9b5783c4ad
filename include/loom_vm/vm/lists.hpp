// -*- C++ -*-
/// @file lists.hpp
/// @brief Type-list and integer-list operations that kernel libraries
/// written for the reference front-end use to build register arrays and
/// statement lists (names and meanings as in the reference's
/// common/tmp.hpp:760-880 and 1676-1760, e.g. vm/lib/linal.hpp:33-130).
///
/// `loom::typelist` works on `type_sequence`; `loom::intlist` on
/// `char_sequence` (register and label names) and `std::integer_sequence`.
#pragma once

#include <cstddef>
#include <type_traits>
#include <utility>

namespace loom
{
template <char... Cs>
struct char_sequence;

template <typename... Ts>
struct type_sequence;

namespace typelist
{
namespace detail
{
template <typename T>
struct is_list : std::false_type
{
};

template <typename... Ts>
struct is_list<type_sequence<Ts...>> : std::true_type
{
};

template <std::size_t I, typename L>
struct at_impl
{
    using type = void; // past the end
};

template <typename T, typename... Ts>
struct at_impl<0, type_sequence<T, Ts...>>
{
    using type = T;
};

template <std::size_t I, typename T, typename... Ts>
    requires(I > 0)
struct at_impl<I, type_sequence<T, Ts...>> : at_impl<I - 1, type_sequence<Ts...>>
{
};

template <typename A, typename B>
struct join_impl;

template <typename... A, typename... B>
struct join_impl<type_sequence<A...>, type_sequence<B...>>
{
    using type = type_sequence<A..., B...>;
};

template <template <typename> typename F, typename L>
struct map_impl;

template <template <typename> typename F, typename... Ts>
struct map_impl<F, type_sequence<Ts...>>
{
    using type = type_sequence<F<Ts>...>;
};

template <typename L>
struct concat_impl;

template <>
struct concat_impl<type_sequence<>>
{
    using type = type_sequence<>;
};

template <typename L, typename... Ls>
struct concat_impl<type_sequence<L, Ls...>>
{
    using type = typename join_impl<L, typename concat_impl<type_sequence<Ls...>>::type>::type;
};

template <typename L>
struct tail_impl;

template <typename T, typename... Ts>
struct tail_impl<type_sequence<T, Ts...>>
{
    using type = type_sequence<Ts...>;
};

template <typename L>
struct length_impl;

template <typename... Ts>
struct length_impl<type_sequence<Ts...>> : std::integral_constant<std::size_t, sizeof...(Ts)>
{
};
} // namespace detail

/// @brief A type_sequence.
template <typename T>
concept List = detail::is_list<T>::value;

/// @brief Number of types in @p L.
template <List L>
inline constexpr std::size_t length = detail::length_impl<L>::value;

/// @brief The @p I-th type of @p L (void past the end).
template <std::size_t I, List L>
using at = typename detail::at_impl<I, L>::type;

/// @brief at<I, T> when @p T is a list, @p T itself otherwise (a scalar
/// operand beside vector operands).
template <std::size_t I, typename T>
using proj = typename std::conditional_t<List<T>, detail::at_impl<I, T>, std::type_identity<T>>::type;

template <List L>
using head = at<0, L>;

template <List L>
using tail = typename detail::tail_impl<L>::type;

/// @brief Concatenation of two lists.
template <List A, List B>
using join = typename detail::join_impl<A, B>::type;

/// @brief @p T in front of @p L.
template <typename T, List L>
using cons = join<type_sequence<T>, L>;

/// @brief @p T after @p L.
template <typename T, List L>
using append = join<L, type_sequence<T>>;

/// @brief F<T> for every T of @p L.
template <template <typename> typename F, List L>
using map = typename detail::map_impl<F, L>::type;

/// @brief Concatenation of a list of lists.
template <List L>
using concat = typename detail::concat_impl<L>::type;
} // namespace typelist

namespace intlist
{
namespace detail
{
template <typename L>
struct traits;

template <char... Cs>
struct traits<char_sequence<Cs...>>
{
    using value_type = char;
    template <char... Xs>
    using make = char_sequence<Xs...>;
    static constexpr std::size_t size = sizeof...(Cs);
};

template <typename T, T... Vs>
struct traits<std::integer_sequence<T, Vs...>>
{
    using value_type = T;
    template <T... Xs>
    using make = std::integer_sequence<T, Xs...>;
    static constexpr std::size_t size = sizeof...(Vs);
};

template <typename L, auto V>
struct append_impl;

template <char... Cs, auto V>
struct append_impl<char_sequence<Cs...>, V>
{
    using type = char_sequence<Cs..., static_cast<char>(V)>;
};

template <typename T, T... Vs, auto V>
struct append_impl<std::integer_sequence<T, Vs...>, V>
{
    using type = std::integer_sequence<T, Vs..., static_cast<T>(V)>;
};

template <typename L, auto V>
struct cons_impl;

template <char... Cs, auto V>
struct cons_impl<char_sequence<Cs...>, V>
{
    using type = char_sequence<static_cast<char>(V), Cs...>;
};

template <typename T, T... Vs, auto V>
struct cons_impl<std::integer_sequence<T, Vs...>, V>
{
    using type = std::integer_sequence<T, static_cast<T>(V), Vs...>;
};

template <typename L>
struct tail_impl;

template <char C, char... Cs>
struct tail_impl<char_sequence<C, Cs...>>
{
    using type = char_sequence<Cs...>;
};

template <typename T, T V, T... Vs>
struct tail_impl<std::integer_sequence<T, V, Vs...>>
{
    using type = std::integer_sequence<T, Vs...>;
};

template <typename L, typename Acc>
struct reverse_impl;

template <typename Acc>
struct reverse_impl<char_sequence<>, Acc>
{
    using type = Acc;
};

template <char C, char... Cs, typename Acc>
struct reverse_impl<char_sequence<C, Cs...>, Acc>
    : reverse_impl<char_sequence<Cs...>, typename cons_impl<Acc, C>::type>
{
};

template <typename T, typename Acc>
struct reverse_impl<std::integer_sequence<T>, Acc>
{
    using type = Acc;
};

template <typename T, T V, T... Vs, typename Acc>
struct reverse_impl<std::integer_sequence<T, V, Vs...>, Acc>
    : reverse_impl<std::integer_sequence<T, Vs...>, typename cons_impl<Acc, V>::type>
{
};

template <typename L>
struct empty_of;

template <char... Cs>
struct empty_of<char_sequence<Cs...>>
{
    using type = char_sequence<>;
};

template <typename T, T... Vs>
struct empty_of<std::integer_sequence<T, Vs...>>
{
    using type = std::integer_sequence<T>;
};

template <template <auto> typename F, typename L>
struct map_t_impl;

template <template <auto> typename F, char... Cs>
struct map_t_impl<F, char_sequence<Cs...>>
{
    using type = type_sequence<F<Cs>...>;
};

template <template <auto> typename F, typename T, T... Vs>
struct map_t_impl<F, std::integer_sequence<T, Vs...>>
{
    using type = type_sequence<F<Vs>...>;
};

template <typename A, typename B>
struct join_impl;

template <char... A, char... B>
struct join_impl<char_sequence<A...>, char_sequence<B...>>
{
    using type = char_sequence<A..., B...>;
};

template <typename T, T... A, T... B>
struct join_impl<std::integer_sequence<T, A...>, std::integer_sequence<T, B...>>
{
    using type = std::integer_sequence<T, A..., B...>;
};
} // namespace detail

/// @brief A char_sequence or std::integer_sequence.
template <typename T>
concept List = requires { detail::traits<T>::size; };

/// @brief Number of values in @p L.
template <List L>
inline constexpr std::size_t length = detail::traits<L>::size;

/// @brief Concatenation of two lists of the same kind.
template <typename A, typename B>
using join = typename detail::join_impl<A, B>::type;

/// @brief @p L followed by @p V.
template <List L, auto V>
using append = typename detail::append_impl<L, V>::type;

/// @brief @p V followed by @p L.
template <auto V, List L>
using cons = typename detail::cons_impl<L, V>::type;

/// @brief @p L without its first value.
template <List L>
using tail = typename detail::tail_impl<L>::type;

/// @brief The values of @p L in reverse order.
template <List L>
using reverse = typename detail::reverse_impl<L, typename detail::empty_of<L>::type>::type;

/// @brief type_sequence<F<V>...> over the values V of @p L.
template <template <auto> typename F, List L>
using map_t = typename detail::map_t_impl<F, L>::type;
} // namespace intlist
} // namespace loom
