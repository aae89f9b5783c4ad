// -*- C++ -*-
/// @file ref_pt2_const.hip
/// @brief TEST INFRASTRUCTURE: the reference's pt2_const kernel lowered for
/// gfx950 by this repository's backend and registered with the library.
#include "registry.hpp"

#include "ref_kernels.hpp"

// The pass kernel needs 138 VGPRs left to itself (three waves per SIMD);
// held to 128 it spills nothing and runs four waves per SIMD.  With the
// regroup paths (239 VGPRs) it is left to itself: held to 128 it spills 324
// bytes per lane and the queue-regrouping run of the 1024x1024 frame takes
// 36 s instead of 24.
LOOM_HIP_WAVES_HINT(4, loom::vm::pt2_const<float, loom_ref::c_b5_t0005_d20>);

#define LOOM_REGISTER_ONE(NAME, ...) LOOM_HIP_REGISTER(NAME, __VA_ARGS__);
LOOM_REF_CATALOG_PT2_CONST(LOOM_REGISTER_ONE)
