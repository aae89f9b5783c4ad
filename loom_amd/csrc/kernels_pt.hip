// -*- C++ -*-
/// @file kernels_pt.hip
/// @brief The built-in path-tracing kernels lowered for gfx950.
#include "registry.hpp"

#include "../kernels/catalog.hpp"

// The path tracer's pass kernel needs 132 VGPRs left to itself (three waves
// per SIMD).  Held to 128 (four waves, two values spilled) it renders the
// 1024x768x64 frame in 8.30 ms instead of 9.05; held to 96 (five waves, 32
// bytes spilled) in 8.11 ms; held to 80 (six waves, 76 bytes) in 8.24 ms.
LOOM_HIP_WAVES_HINT(5, loom_amd::smallpt::smallpt9<float>);
// With the regroup paths it needs 133: held to 128 (four waves) a
// regrouping frame with refill takes 8.5 ms, held to 96 (five waves, 40
// bytes spilled) 8.3 ms, and the first pass of a regrouping frame without
// refill 16.2-16.6 ms instead of 18.0-18.4 left to itself.
LOOM_HIP_REGROUP_WAVES_HINT(5, loom_amd::smallpt::smallpt9<float>);

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_AMD_CATALOG_PT(LOOM_REGISTER_ONE)
