// -*- C++ -*-
/// @file ref_pt2_const_1024.hip
/// @brief TEST INFRASTRUCTURE: the reference's pt2_const kernel (1024x1024 frame) lowered for
/// gfx950 by this repository's backend and registered with the library.
#include "registry.hpp"

#include "ref_kernels.hpp"

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_REF_CATALOG_PT2_CONST_1024(LOOM_REGISTER_ONE)
