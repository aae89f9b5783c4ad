// -*- C++ -*-
/// @file catalog.hpp
/// @brief The built-in kernels: one list, expanded by the gfx950 library
/// (registration) and by the CPU oracle, so both address kernels by the same
/// names.  X(name, kernel type).
#pragma once

#include "extras.hpp"
#include "programs.hpp"
#include "smallpt9.hpp"

#define LOOM_AMD_CATALOG_CORE(X)                                      \
    X("fib_u64", loom_amd::programs::fib<std::uint64_t>)              \
    X("fib_u32", loom_amd::programs::fib<std::uint32_t>)              \
    X("collatz_u64", loom_amd::programs::collatz<std::uint64_t>)      \
    X("factorial_u64", loom_amd::programs::factorial<std::uint64_t>)  \
    X("jump4_u32", loom_amd::programs::jump4<std::uint32_t>)          \
    X("int_mix_u64", loom_amd::programs::int_mix<std::uint64_t>)      \
    X("int_mix_u32", loom_amd::programs::int_mix<std::uint32_t>)      \
    X("fp_mix_f32", loom_amd::programs::fp_mix<float>)                \
    X("fp_mix_f64", loom_amd::programs::fp_mix<double>)               \
    X("sincos_f32", loom_amd::programs::sincos<float>)                \
    X("sincos_f64", loom_amd::programs::sincos<double>)               \
    X("roots_f32", loom_amd::programs::roots<float>)                  \
    X("roots_f64", loom_amd::programs::roots<double>)                 \
    X("literal_div_f32", loom_amd::programs::literal_div<float>)      \
    X("literal_div_f64", loom_amd::programs::literal_div<double>)     \
    X("sct_range_f32", loom_amd::programs::sct_range<float>)          \
    X("sct_range_f64", loom_amd::programs::sct_range<double>)         \
    X("threefry2x32_u32", loom_amd::programs::threefry2x32<std::uint32_t>) \
    X("cubic_bisect_f32", loom_amd::programs::cubic_bisect<float>)    \
    X("cubic_bisect_f64", loom_amd::programs::cubic_bisect<double>)

#define LOOM_AMD_CATALOG_PT(X) X("smallpt9_f32", loom_amd::smallpt::smallpt9<float>)

#define LOOM_AMD_CATALOG(X) LOOM_AMD_CATALOG_CORE(X) LOOM_AMD_CATALOG_PT(X)
