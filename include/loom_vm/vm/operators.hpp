// -*- C++ -*-
/// @file operators.hpp
/// @brief Operation tags used in `vm::Compute`, source-compatible with the
/// reference (src/vm/operators.hpp:81-330).
///
/// Besides these tags, `Compute` accepts the standard functors `std::plus`,
/// `std::minus`, `std::multiplies`, `std::divides`, `std::modulus`,
/// `std::negate`, the six comparisons, `std::logical_and/or/not`,
/// `std::bit_and/or/xor/not` and `std::identity` (move with conversion).
///
/// The tags carry no arithmetic here: the op contract (exact semantics shared
/// by the gfx950 lowering and the CPU oracle) is implemented by the device
/// lowering in `loom_amd/csrc/device/ops.hpp`.
#pragma once

#include <vm/core.hpp>

#include <functional>

namespace loom::vm::ops
{
/// @brief `c ? a : b` (operators.hpp:86-96).
struct cmov
{
};

/// @brief `v << s` (operators.hpp:101-110).  Shift counts are taken modulo
/// the bit width.
template <typename T>
struct lshift
{
};

/// @brief `v >> s` (operators.hpp:115-126), arithmetic for signed types.
template <typename T>
struct rshift
{
};

/// @brief |v| (operators.hpp:129-137).
template <typename T>
struct abs
{
};

/// @brief Fused `x * y + z` (operators.hpp:142-150).
struct fma
{
};

/// @brief Fused `x * y - z` (operators.hpp:155-163).
struct fms
{
};

/// @brief `1 / sqrt(x)` (operators.hpp:166-174).
template <typename T>
struct rsqrt
{
};

/// @brief Select a constant from the list @p L by index (operators.hpp:183-202).
template <typename T, typename I, typename L>
    requires(L::size >= 3)
struct sct
{
    using type  = T;
    using index = I;
    using list  = L;
};

/// @brief Select a register by index (operators.hpp:209-219).
template <typename T, typename I>
struct slt
{
};

/// @brief Predicate accumulation `Acc(p, Cmp(x, y))` (operators.hpp:239-254).
template <typename T,
          typename I,
          template <typename> typename Acc,
          template <typename> typename Cmp>
struct pacc
{
};

/// @brief Checks that a floating-point value is finite (operators.hpp:256-276).
/// On the GPU a failure sets the run's error flag instead of aborting.
struct assert_fp
{
};

template <typename T>
struct max
{
};
template <typename T>
struct min
{
};
template <typename T>
struct round
{
};
template <typename T>
struct ceil
{
};
template <typename T>
struct floor
{
};
template <typename T>
struct trunc
{
};
template <typename T>
struct sin
{
};
template <typename T>
struct cos
{
};
template <typename T>
struct sqrt
{
};
template <typename T>
struct copysign
{
};
} // namespace loom::vm::ops
