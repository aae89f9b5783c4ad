// -*- C++ -*-
/// @file kernels_core.hip
/// @brief Built-in integer and floating-point kernels lowered for gfx950.
#include "registry.hpp"

#include "../kernels/catalog.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_AMD_CATALOG_CORE(LOOM_REGISTER_ONE)
