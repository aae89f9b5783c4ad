// -*- C++ -*-
/// @file kernels_pt.hip
/// @brief The built-in path-tracing kernels lowered for gfx950.
#include "registry.hpp"

#include "../kernels/catalog.hpp"

// The path tracer's pass kernel needs 132 VGPRs left to itself (three waves
// per SIMD); held to 128 (four waves, two values spilled) it renders the
// 1024x768x64 frame about 10% faster.
LOOM_HIP_WAVES_HINT(4, loom_amd::smallpt::smallpt9<float>);

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_AMD_CATALOG_PT(LOOM_REGISTER_ONE)
