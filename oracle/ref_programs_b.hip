// -*- C++ -*-
/// @file ref_programs_b.hip
/// @brief TEST INFRASTRUCTURE: one group of the reference's test programs
/// (src/tests/programs.hpp) lowered for gfx950 by this repository's backend
/// and registered with the library.
#include "registry.hpp"

#include "ref_kernels.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_REF_CATALOG_PROGRAMS_B(LOOM_REGISTER_ONE)
