// -*- C++ -*-
/// @file ref_validators.cpp
/// @brief ORACLE / TEST INFRASTRUCTURE: the reference's own validation
/// functions for its test programs (src/tests/programs.hpp:288-391, fn::val_*,
/// and the std functors src/tests/vm.cpp:300-311 pairs with the bit and
/// boolean programs), compiled from the reference checkout into
/// oracle/_ref/liboracle_ref.so beside the interpreter.  A test applies the
/// validator named after a program to the same input columns and compares
/// with the outputs of the interpreter and of the GPU backend, as the
/// reference's suites compare their engines with these functions
/// (src/tests/vm.cpp:241-246, vm.cu:58-82).
///
/// fp_mix_k (src/benchmarks/fp_mix.hpp) has its reference fp_mix_ref, which
/// the reference compares within 100 ULP on 99% of inputs
/// (src/tests/vm_test_suite_utils.hpp:150-182); ref_ulp_mismatches applies
/// the reference's almost_equal (src/common/math.hpp:253-274) for that.
///
/// ref_rng applies the reference's C++ Threefry-2x32, bump_ctr and
/// gen_canonical (src/common/random.hpp:37-103), the functions its own test
/// holds the IR macros to (src/tests/vm_smallpt.cpp:358-376).
///
/// The programs checked by an invariant instead (abs_test ... rsqrt_test,
/// sct_inv_test, slt_inv_test: every output non-zero,
/// src/common/test.hpp:636-648) have no validator here.
#include <common/math.hpp> // before fp_mix.hpp, whose fp_mix_ref calls loom::fma
#include <benchmarks/fp_mix.hpp>
#include <common/random.hpp>
#include <tests/programs.hpp>

#include <cstdint>
#include <cstring>
#include <functional>
#include <utility>

namespace
{
using Fn = void (*)(const void* const* in, void* out, std::uint64_t n);

template <typename Out, typename... In, typename V, std::size_t... I>
void
apply(V v, const void* const* in, void* out, std::uint64_t n, std::index_sequence<I...>)
{
    Out* o = static_cast<Out*>(out);
    for(std::uint64_t t = 0; t < n; ++t)
        o[t] = static_cast<Out>(v(static_cast<const In*>(in[I])[t]...));
}

template <typename Out, typename... In, typename V>
void
apply(V v, const void* const* in, void* out, std::uint64_t n)
{
    apply<Out, In...>(v, in, out, n, std::index_sequence_for<In...>{});
}

struct Entry
{
    const char* name;
    Fn          fn;
};

// X(program name, output type, validator object, input types...)
#define REF_V(NAME, OUT, V, ...) \
    {NAME, [](const void* const* in, void* out, std::uint64_t n) { apply<OUT, __VA_ARGS__>(V, in, out, n); }},
#define REF_V_INT(NAME, V)                                                                \
    REF_V("ref_" NAME "_i8", std::int8_t, V, std::int8_t)                                 \
    REF_V("ref_" NAME "_i16", std::int16_t, V, std::int16_t)                              \
    REF_V("ref_" NAME "_i32", std::int32_t, V, std::int32_t)                              \
    REF_V("ref_" NAME "_i64", std::int64_t, V, std::int64_t)
#define REF_V_UINT1(NAME, V)                                                              \
    REF_V("ref_" NAME "_u8", std::uint8_t, V, std::uint8_t)                               \
    REF_V("ref_" NAME "_u16", std::uint16_t, V, std::uint16_t)                            \
    REF_V("ref_" NAME "_u32", std::uint32_t, V, std::uint32_t)                            \
    REF_V("ref_" NAME "_u64", std::uint64_t, V, std::uint64_t)
#define REF_V_UINT2(NAME, V)                                                              \
    REF_V("ref_" NAME "_u8", std::uint8_t, V, std::uint8_t, std::uint8_t)                 \
    REF_V("ref_" NAME "_u16", std::uint16_t, V, std::uint16_t, std::uint16_t)             \
    REF_V("ref_" NAME "_u32", std::uint32_t, V, std::uint32_t, std::uint32_t)             \
    REF_V("ref_" NAME "_u64", std::uint64_t, V, std::uint64_t, std::uint64_t)

const Entry kEntries[] = {
    REF_V_INT("increment", val_inc)
    REF_V_UINT1("increment", val_inc)
    REF_V("ref_fac_test_u32", std::uint32_t, val_fac, std::uint32_t)
    REF_V("ref_fac_test_u64", std::uint64_t, val_fac, std::uint64_t)
    REF_V("ref_fib_test_u64", std::uint64_t, val_fib, std::uint64_t)
    REF_V("ref_union_test_f32", std::uint32_t, val_union1, float)
    REF_V("ref_union_test_f64", std::uint64_t, val_union1, double)
    REF_V("ref_union_cmov_test_f32", std::uint32_t, val_union2, float)
    REF_V("ref_union_cmov_test_f64", std::uint64_t, val_union2, double)
    REF_V_UINT1("jmp4_test", val_jmp4)
    REF_V("ref_jmp4_test_f32", std::uint32_t, val_jmp4, std::uint32_t)
    REF_V("ref_lshift_test_u16", std::uint16_t, val_lshift, std::uint16_t, std::uint16_t)
    REF_V("ref_lshift_test_u32", std::uint32_t, val_lshift, std::uint32_t, std::uint32_t)
    REF_V("ref_lshift_test_u64", std::uint64_t, val_lshift, std::uint64_t, std::uint64_t)
    REF_V("ref_rshift_test_u16", std::uint16_t, val_rshift, std::uint16_t, std::uint16_t)
    REF_V("ref_rshift_test_u32", std::uint32_t, val_rshift, std::uint32_t, std::uint32_t)
    REF_V("ref_rshift_test_u64", std::uint64_t, val_rshift, std::uint64_t, std::uint64_t)
    REF_V_UINT2("bit_and_test", std::bit_and<>{})
    REF_V_UINT2("bit_or_test", std::bit_or<>{})
    REF_V_UINT2("bit_xor_test", std::bit_xor<>{})
    REF_V_UINT1("bit_not_test", val_bit_not)
    REF_V_UINT2("bool_and_test", std::logical_and<>{})
    REF_V_UINT2("bool_or_test", std::logical_or<>{})
    REF_V_UINT1("bool_not_test", std::logical_not<>{})
    REF_V("ref_fp_mix_k_f32", float, fp_mix_ref{}, float, float, float)
    REF_V("ref_fp_mix_k_f64", double, fp_mix_ref{}, double, double, double)
};

template <typename T>
std::uint64_t
ulp_mismatches(const T* a, const T* b, std::uint64_t n, unsigned ulp)
{
    std::uint64_t bad = 0;
    for(std::uint64_t t = 0; t < n; ++t) bad += !loom::almost_equal(a[t], b[t], ulp);
    return bad;
}
} // namespace

/// Applies the reference's validator for program @p name to @p n threads of
/// the input columns @p in, writing the expected output column @p out.
/// Returns 0, or -1 when the program has no validator.
extern "C" int
ref_validate(const char* name, const void* const* in, void* out, std::uint64_t n)
{
    for(const Entry& e : kEntries)
        if(std::strcmp(e.name, name) == 0)
        {
            e.fn(in, out, n);
            return 0;
        }
    return -1;
}

/// Number of the @p n pairs a[t], b[t] that the reference's
/// loom::almost_equal(a, b, ulp) rejects; @p width 4 for float, 8 for double.
extern "C" std::uint64_t
ref_ulp_mismatches(const void* a, const void* b, std::uint64_t n, unsigned ulp, int width)
{
    if(width == 4)
        return ulp_mismatches(static_cast<const float*>(a), static_cast<const float*>(b), n, ulp);
    return ulp_mismatches(static_cast<const double*>(a), static_cast<const double*>(b), n, ulp);
}

/// The reference's C++ random-number functions on counters (c0[t], c1[t]):
/// the Threefry-2x32 block under key (0xdead, 0xbeef) into (w0, w1), the
/// counter bumped by 39916801 into (n0, n1), and gen_canonical of the block
/// words into (f0, f1).
extern "C" void
ref_rng(const std::uint32_t* c0, const std::uint32_t* c1, std::uint32_t* w0, std::uint32_t* w1,
        std::uint32_t* n0, std::uint32_t* n1, float* f0, float* f1, std::uint64_t n)
{
    for(std::uint64_t t = 0; t < n; ++t)
    {
        const loom::fn::ar_32_2 c{c0[t], c1[t]};
        const loom::fn::ar_32_2 w = loom::threefry2x32r32<0xdead, 0xbeef>(c);
        const loom::fn::ar_32_2 b = loom::bump_ctr(c, 39916801u);
        w0[t]                     = w[0];
        w1[t]                     = w[1];
        n0[t]                     = b[0];
        n1[t]                     = b[1];
        f0[t]                     = loom::gen_canonical(w[0]);
        f1[t]                     = loom::gen_canonical(w[1]);
    }
}
