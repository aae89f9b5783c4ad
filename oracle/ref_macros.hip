// -*- C++ -*-
/// @file ref_macros.hip
/// @brief TEST INFRASTRUCTURE: the kernels applying the reference's
/// path-tracing and vector IR macros one at a time (ref_macros.hpp),
/// lowered for gfx950 by this repository's backend and registered with the
/// library.
#include "registry.hpp"

#include "ref_kernels.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_REF_CATALOG_MACROS(LOOM_REGISTER_ONE)
