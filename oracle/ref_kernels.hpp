// -*- C++ -*-
/// @file ref_kernels.hpp
/// @brief ORACLE / TEST INFRASTRUCTURE: the reference's own path-tracing
/// kernels (src/smallpt/vm/pt.hpp:1259-1367 pt_const, 1424-1729 pt2_const),
/// included from the reference checkout and lowered by this repository's
/// front-end — the drop-in check that existing kernels run unchanged.  The
/// configurations are two of the reference's benchmarked ones
/// (src/benchmarks/pt.cpp:31-35, CMakeLists.txt:419-460): 40x40 pixels, 64
/// samples at least, blur / threshold / maximum depth as in the names; and
/// the second one on a 1024x1024 frame.
/// Built only by oracle/build_ref.py, into oracle/_ref/, where the reference
/// sources exist.
#pragma once

#include <smallpt/vm/pt.hpp>

namespace loom_ref
{
using c_b1_t0001_d2  = loom::vm::constants<CS(40), loom::sl2cs<"0.0001">, CS(64),
                                           loom::sl2cs<"1">, loom::sl2cs<"2">>;
using c_b5_t0005_d20 = loom::vm::constants<CS(40), loom::sl2cs<"0.0005">, CS(64),
                                           loom::sl2cs<"5">, loom::sl2cs<"20">>;
/// the same configuration on a 1024x1024 frame (a GPU-sized launch)
using c_b5_t0005_d20_1024 = loom::vm::constants<CS(1024), loom::sl2cs<"0.0005">, CS(64),
                                                loom::sl2cs<"5">, loom::sl2cs<"20">>;
} // namespace loom_ref

/// X(name, kernel type) for every reference kernel of the check.
#define LOOM_REF_CATALOG_PT_CONST(X) \
    X("ref_pt_const_b1_t0.0001_d2", loom::vm::pt_const<float, loom_ref::c_b1_t0001_d2>)
#define LOOM_REF_CATALOG_PT2_CONST(X) \
    X("ref_pt2_const_b5_t0.0005_d20", loom::vm::pt2_const<float, loom_ref::c_b5_t0005_d20>)
#define LOOM_REF_CATALOG_PT2_CONST_1024(X) \
    X("ref_pt2_const_b5_t0.0005_d20_1024", \
      loom::vm::pt2_const<float, loom_ref::c_b5_t0005_d20_1024>)
#define LOOM_REF_CATALOG(X) \
    LOOM_REF_CATALOG_PT_CONST(X) LOOM_REF_CATALOG_PT2_CONST(X) LOOM_REF_CATALOG_PT2_CONST_1024(X)
