// -*- C++ -*-
/// @file ref_kernels.hpp
/// @brief ORACLE / TEST INFRASTRUCTURE: the reference's own path-tracing
/// kernels (src/smallpt/vm/pt.hpp:1259-1367 pt_const, 1424-1729 pt2_const),
/// included from the reference checkout and lowered by this repository's
/// front-end — the drop-in check that existing kernels run unchanged.  The
/// configurations are two of the reference's benchmarked ones
/// (src/benchmarks/pt.cpp:31-35, CMakeLists.txt:419-460): 40x40 pixels, 64
/// samples at least, blur / threshold / maximum depth as in the names; and
/// the second one on a 1024x1024 frame.
/// Also the reference's own test programs (src/tests/programs.hpp), at the
/// element types its VM test suites instantiate them with
/// (src/tests/vm.cpp:262-325 for the scalar and vector engines,
/// src/tests/vm_simd.cpp:71-90 and vm_simd_compact.cpp:180-251 for the SIMD
/// and compact engines, src/tests/vm.cu:31-41 for the CUDA backend), and
/// the floating-point benchmark kernel fp_mix_k (src/benchmarks/fp_mix.hpp:
/// 95-119; tested by src/tests/vm_fp_mix.cpp:33-41).
/// Built only by oracle/build_ref.py, into oracle/_ref/, where the reference
/// sources exist.
#pragma once

// common/math.hpp first: it declares loom::fac and loom::fma, which the
// validators in programs.hpp and fp_mix.hpp call
#include <common/math.hpp>
#include <benchmarks/fp_mix.hpp>
#include <smallpt/vm/pt.hpp>
#include <tests/programs.hpp>

#include "ref_macros.hpp"

namespace loom_ref
{
using c_b1_t0001_d2  = loom::vm::constants<CS(40), loom::sl2cs<"0.0001">, CS(64),
                                           loom::sl2cs<"1">, loom::sl2cs<"2">>;
using c_b5_t0005_d20 = loom::vm::constants<CS(40), loom::sl2cs<"0.0005">, CS(64),
                                           loom::sl2cs<"5">, loom::sl2cs<"20">>;
/// the same configuration on a 1024x1024 frame (a GPU-sized launch)
using c_b5_t0005_d20_1024 = loom::vm::constants<CS(1024), loom::sl2cs<"0.0005">, CS(64),
                                                loom::sl2cs<"5">, loom::sl2cs<"20">>;
} // namespace loom_ref

namespace loom_ref
{
// clang-format off
/// The reference's random-number IR macros (src/vm/lib/random.hpp:30-85) in
/// the order its path tracer applies them (src/smallpt/vm/pt.hpp:890-895,
/// 277-278): the Threefry-2x32 block of counter (c0, c1) under the key its
/// benchmarks use (0xdead, 0xbeef), the counter bumped by 39916801, and both
/// block words as canonical floats.  Outputs: the block words, the bumped
/// counter, the two floats.
template <std::floating_point T>
KERNEL(rng_macros, IVAR(c0), IVAR(c1))
    IMOV(u0, c0),
    IMOV(u1, c1),
    INSERT_MACRO(loom::vm::threefry2x32r32, CS(u0), CS(u1), CS(0xdead), CS(0xbeef), CS(0x1bd17b98),
                 CS(tmp0), CS(tmp1)),
    INSERT_MACRO(loom::vm::bump_ctr, CS(c0), CS(c1), CS(39916801), CS(tmp0), CS(tmp1), CS(tmp2)),
    IMOV(w0, u0),
    IMOV(w1, u1),
    INSERT_TMACRO(loom::vm::gen_canonical, CS(u0)),
    INSERT_TMACRO(loom::vm::gen_canonical, CS(u1))
KERNEL_END(IVAR(w0), IVAR(w1), IVAR(c0), IVAR(c1), TVAR(u0), TVAR(u1));
// clang-format on
} // namespace loom_ref

/// X(name, kernel type) for every reference kernel of the check.
#define LOOM_REF_CATALOG_PT_CONST(X) \
    X("ref_pt_const_b1_t0.0001_d2", loom::vm::pt_const<float, loom_ref::c_b1_t0001_d2>)
#define LOOM_REF_CATALOG_PT2_CONST(X) \
    X("ref_pt2_const_b5_t0.0005_d20", loom::vm::pt2_const<float, loom_ref::c_b5_t0005_d20>)
#define LOOM_REF_CATALOG_PT2_CONST_1024(X) \
    X("ref_pt2_const_b5_t0.0005_d20_1024", \
      loom::vm::pt2_const<float, loom_ref::c_b5_t0005_d20_1024>)

/// The test programs, grouped so that each group is one compilation unit.
#define LOOM_REF_EACH_INT(X, NAME, K)                                          \
    X("ref_" NAME "_i8", K<std::int8_t>) X("ref_" NAME "_i16", K<std::int16_t>) \
    X("ref_" NAME "_i32", K<std::int32_t>) X("ref_" NAME "_i64", K<std::int64_t>)
#define LOOM_REF_EACH_UINT(X, NAME, K)                                           \
    X("ref_" NAME "_u8", K<std::uint8_t>) X("ref_" NAME "_u16", K<std::uint16_t>) \
    X("ref_" NAME "_u32", K<std::uint32_t>) X("ref_" NAME "_u64", K<std::uint64_t>)
#define LOOM_REF_EACH_FLOAT(X, NAME, K) \
    X("ref_" NAME "_f32", K<float>) X("ref_" NAME "_f64", K<double>)

#define LOOM_REF_CATALOG_PROGRAMS_A(X)                                     \
    LOOM_REF_EACH_INT(X, "increment", increment)                           \
    LOOM_REF_EACH_UINT(X, "increment", increment)                          \
    X("ref_fac_test_u32", fac_test<std::uint32_t>)                         \
    X("ref_fac_test_u64", fac_test<std::uint64_t>)                         \
    X("ref_fib_test_u64", fib_test<std::uint64_t>)                         \
    LOOM_REF_EACH_FLOAT(X, "union_test", union_test)                       \
    LOOM_REF_EACH_FLOAT(X, "union_cmov_test", union_cmov_test)             \
    LOOM_REF_EACH_UINT(X, "jmp4_test", jmp4_test)                          \
    X("ref_jmp4_test_f32", jmp4_test<float>)                               \
    X("ref_lshift_test_u16", lshift_test<std::uint16_t>)                   \
    X("ref_lshift_test_u32", lshift_test<std::uint32_t>)                   \
    X("ref_lshift_test_u64", lshift_test<std::uint64_t>)                   \
    X("ref_rshift_test_u16", rshift_test<std::uint16_t>)                   \
    X("ref_rshift_test_u32", rshift_test<std::uint32_t>)                   \
    X("ref_rshift_test_u64", rshift_test<std::uint64_t>)
#define LOOM_REF_CATALOG_PROGRAMS_B(X)                                     \
    LOOM_REF_EACH_UINT(X, "bit_and_test", bit_and_test)                    \
    LOOM_REF_EACH_UINT(X, "bit_or_test", bit_or_test)                      \
    LOOM_REF_EACH_UINT(X, "bit_xor_test", bit_xor_test)                    \
    LOOM_REF_EACH_UINT(X, "bit_not_test", bit_not_test)                    \
    LOOM_REF_EACH_UINT(X, "bool_and_test", bool_and_test)                  \
    LOOM_REF_EACH_UINT(X, "bool_or_test", bool_or_test)                    \
    LOOM_REF_EACH_UINT(X, "bool_not_test", bool_not_test)                  \
    LOOM_REF_EACH_UINT(X, "sct_inv_test", sct_inv_test)                    \
    LOOM_REF_EACH_UINT(X, "slt_inv_test", slt_inv_test)                    \
    X("ref_rng_macros_f32", loom_ref::rng_macros<float>)
#define LOOM_REF_CATALOG_PROGRAMS_C(X)                                     \
    LOOM_REF_EACH_FLOAT(X, "abs_test", abs_test)                           \
    LOOM_REF_EACH_FLOAT(X, "trunc_test", trunc_test)                       \
    LOOM_REF_EACH_FLOAT(X, "ceilfloor", ceilfloor)                         \
    LOOM_REF_EACH_FLOAT(X, "minmax_test", minmax_test)                     \
    LOOM_REF_EACH_FLOAT(X, "arithmetic_test", arithmetic_test)             \
    LOOM_REF_EACH_FLOAT(X, "sin_cos_test", sin_cos_test)                   \
    LOOM_REF_EACH_FLOAT(X, "sqrt_test", sqrt_test)                         \
    LOOM_REF_EACH_FLOAT(X, "rsqrt_test", rsqrt_test)                       \
    LOOM_REF_EACH_FLOAT(X, "fp_mix_k", fp_mix_k)
#define LOOM_REF_CATALOG_PROGRAMS(X) \
    LOOM_REF_CATALOG_PROGRAMS_A(X) LOOM_REF_CATALOG_PROGRAMS_B(X) LOOM_REF_CATALOG_PROGRAMS_C(X)

#define LOOM_REF_CATALOG(X)                                                                       \
    LOOM_REF_CATALOG_PT_CONST(X) LOOM_REF_CATALOG_PT2_CONST(X) LOOM_REF_CATALOG_PT2_CONST_1024(X) \
    LOOM_REF_CATALOG_PROGRAMS(X) LOOM_REF_CATALOG_MACROS(X)
