// -*- C++ -*-
/// @file ref_pt2_const.hip
/// @brief TEST INFRASTRUCTURE: the reference's pt2_const kernel lowered for
/// gfx950 by this repository's backend and registered with the library.
#include "registry.hpp"

#include "ref_kernels.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_REF_CATALOG_PT2_CONST(LOOM_REGISTER_ONE)
