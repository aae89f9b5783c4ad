// -*- C++ -*-
/// @file kernels_pt.hip
/// @brief The built-in path-tracing kernels lowered for gfx950.
#include "registry.hpp"

#include "../kernels/catalog.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_AMD_CATALOG_PT(LOOM_REGISTER_ONE)
