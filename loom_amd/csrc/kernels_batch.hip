// -*- C++ -*-
/// @file kernels_batch.hip
/// @brief Batch sets: persistent kernels that mix several built-in kernels.
#include "registry.hpp"

#include "../kernels/catalog.hpp"

namespace P = loom_amd::programs;
namespace S = loom_amd::smallpt;

// the mixed benchmark batch: path-trace frames with fib and Collatz jobs
LOOM_HIP_REGISTER_BATCH("pt_int_mix", S::smallpt9<float>, P::fib<std::uint64_t>,
                        P::collatz<std::uint64_t>);
// integer-only batch (a smaller register file, higher occupancy)
LOOM_HIP_REGISTER_BATCH("int_mix", P::fib<std::uint64_t>, P::collatz<std::uint64_t>,
                        P::factorial<std::uint64_t>);
