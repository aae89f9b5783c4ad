// -*- C++ -*-
/// @file ref_macros.hpp
/// @brief ORACLE / TEST INFRASTRUCTURE: kernels that apply the reference's
/// path-tracing and vector IR macros (src/smallpt/vm/pt.hpp,
/// src/vm/lib/linal.hpp, included from the reference checkout) one at a
/// time, as the reference's own macro checks do (src/tests/vm_smallpt.cpp:
/// 36-227, src/tests/vm_linal.cpp:29-66).  Each kernel's outputs are
/// compared with C++ functions (oracle/ref_validators.cpp): the reference's
/// own C++ where it compiles here (primitive intersections, the quadratic
/// solver, the orthonormal basis, disk sampling, vector algebra) and a
/// restatement of its C++ formula elsewhere; and the GPU's outputs with the
/// oracle's bit for bit.  Built only by oracle/build_ref.py.
#pragma once

#include <smallpt/vm/pt.hpp>
#include <vm/lib/linal.hpp>

namespace loom_ref::pm
{
namespace vm = loom::vm;
// clang-format off

template <std::floating_point T>
TKERNEL13(disk, o0, o1, o2, d0, d1, d2, p0, p1, p2, n0, n1, n2, r)
    INSERT_TMACRO(vm::isect_disk, CS(hit), CS(t), CS(o), CS(d), CS(p), CS(n), CS(r), CS(vtmp), CS(tmp))
KERNEL_END(IVAR(hit), TVAR(t));

template <std::floating_point T>
TKERNEL3(quadratic, a, b, c)
    INSERT_TMACRO(vm::solve_quadratic, CS(s), CS(x), CS(a), CS(b), CS(c), CS(t), CS(t2))
KERNEL_END(IVAR(s), TVAR(x));

template <std::floating_point T>
TKERNEL10(sphere, o0, o1, o2, d0, d1, d2, p0, p1, p2, r)
    INSERT_TMACRO(vm::prep_isect_sphere, CS(a), CS(b), CS(c), CS(o), CS(d), CS(p), CS(r), CS(tmp0), CS(tmp1), CS(vtmp1)),
    INSERT_TMACRO(vm::solve_quadratic, CS(hit), CS(t), CS(a), CS(b), CS(c), CS(tmp0), CS(tmp1))
KERNEL_END(IVAR(hit), TVAR(t));

template <std::floating_point T>
TKERNEL11(cylinder, o0, o1, o2, d0, d1, d2, p0, p1, p2, r, h)
    INSERT_TMACRO(vm::vec_vec3_sub, CS(o), CS(o), CS(p)),
    INSERT_TMACRO(vm::cyl_prep_abc, CS(a), CS(b), CS(c), CS(o0), CS(o1), CS(o2), CS(d0), CS(d1), CS(d2), CS(r)),
    INSERT_TMACRO(vm::solve_quadratic, CS(hit), CS(t), CS(a), CS(b), CS(c), CS(tmp0), CS(tmp1)),
    INSERT_TMACRO(vm::cyl_fin, CS(hit), CS(t), CS(o1), CS(d1), CS(tmp0), CS(h))
KERNEL_END(IVAR(hit), TVAR(t));

template <std::floating_point T>
TKERNEL15(triangle, o0, o1, o2, d0, d1, d2, vert00, vert01, vert02, vert10, vert11, vert12,
          vert20, vert21, vert22)
    INSERT_TMACRO(vm::isect_triangle, CS(hit), CS(t), CS(o), CS(d), CS(vert0), CS(vert1), CS(vert2), CS(tmp1), CS(tmp2), CS(vtmp))
KERNEL_END(IVAR(hit), TVAR(t));

template <std::floating_point T>
TKERNEL18(quad, o0, o1, o2, d0, d1, d2, v000, v001, v002, v010, v011, v012, v110, v111, v112,
          v100, v101, v102)
    INSERT_TMACRO(vm::isect_quad, CS(hit), CS(t), CS(o), CS(d), CS(v00), CS(v01), CS(v11), CS(v10),
                  CS(vt1), CS(vt2), CS(vt3), CS(tmp1), CS(tmp2), CS(tmp3))
KERNEL_END(IVAR(hit), TVAR(t));

/// orthonormal basis of n, both vectors normalised afterwards
template <std::floating_point T>
TKERNEL3(onb, n0, n1, n2)
    INSERT_TMACRO(vm::onb, CS(b0), CS(b1), CS(b2), CS(t0), CS(t1), CS(t2), CS(n0), CS(n1), CS(n2)),
    INSERT_TMACRO(vm::vec3_norm, CS(b), CS(tmp)),
    INSERT_TMACRO(vm::vec3_norm, CS(t), CS(tmp))
TKERNEL_END6(b0, b1, b2, t0, t1, t2);

template <std::floating_point T>
TKERNEL5(hemisphere, n0, n1, n2, u0, u1)
    INSERT_TMACRO(vm::onb, CS(b0), CS(b1), CS(b2), CS(t0), CS(t1), CS(t2), CS(n0), CS(n1), CS(n2)),
    INSERT_TMACRO(vm::vec3_norm, CS(b), CS(tmp0)),
    INSERT_TMACRO(vm::vec3_norm, CS(t), CS(tmp0)),
    INSERT_TMACRO(vm::sample_hemisphere, CS(d), CS(n), CS(b), CS(t), CS(u0), CS(u1), CS(tmp0), CS(tmp1)),
    INSERT_TMACRO(vm::vec3_norm, CS(d), CS(tmp0))
TKERNEL_END3(d0, d1, d2);

/// pixel (x, y) of a w x h film to sample space; integer inputs
template <std::floating_point T>
KERNEL(sample_space, IVAR(x), IVAR(y), IVAR(w), IVAR(h))
    MOV(TVAR(w), IVAR(w)),
    MOV(TVAR(h), IVAR(h)),
    MOV(TVAR(x), IVAR(x)),
    MOV(TVAR(y), IVAR(y)),
    TDIV(r, w, h),
    INSERT_TMACRO(vm::to_sample_space, CS(x), CS(y), CS(w), CS(h), CS(r))
TKERNEL_END2(x, y);

/// camera ray direction of normalised device coordinates (d0, d1)
template <std::floating_point T>
TKERNEL2(ray_dir, d0, d1)
    INSERT_TMACRO(vm::gen_ray_dir, CS(d0), CS(d1)),
    TMOV_CS(CS(d2), vm::scene::CAM_Dz),
    INSERT_TMACRO(vm::vec3_norm, CS(d), CS(tmp))
TKERNEL_END3(d0, d1, d2);

template <std::floating_point T>
TKERNEL2(light, u0, u1)
    INSERT_TMACRO(vm::sample_light_disk, CS(dst), CS(u0), CS(u1), CS(tmp0), CS(vtmp0))
TKERNEL_END3(dst0, dst1, dst2);

template <std::floating_point T>
TKERNEL2(schlick, cd, eta)
    INSERT_TMACRO(vm::schlick, CS(r), CS(cd), CS(eta), CS(tmp0), CS(tmp1))
TKERNEL_END1(r);

/// reflection of direction i about normal n
template <std::floating_point T>
TKERNEL6(reflect, n0, n1, n2, i0, i1, i2)
    INSERT_TMACRO(vm::vec3_negate, CS(wi), CS(i)),
    INSERT_TMACRO(vm::vec3_dot, CS(ct), CS(n), CS(wi)),
    INSERT_TMACRO(vm::reflect, CS(r), CS(n), CS(i), CS(ct))
TKERNEL_END3(r0, r1, r2);

/// refraction of direction d at normal n into glass (index 1.5), the
/// normal flipped to face d and |cos^2 t| as the reference's test forms it
template <std::floating_point T>
TKERNEL6(refract, n0, n1, n2, d0, d1, d2)
    INSERT_TMACRO(vm::comp_f, CS(f), CS(wi), CS(ct), CS(into), CS(eta), CS(1.5), CS(tmp0), CS(tmp1)),
    INSERT_TMACRO(vm::vec3_negate, CS(ng), CS(n)),
    TCMOV(n0, into, n0, ng0),
    TCMOV(n1, into, n1, ng1),
    TCMOV(n2, into, n2, ng2),
    INSERT_TMACRO(vm::vec3_dot, CS(ddn), CS(d), CS(n)),
    TMUL(tmp0, ddn, ddn),
    TSUB(tmp0, 1, tmp0),
    TMUL(tmp0, eta, tmp0),
    TMUL(tmp0, eta, tmp0),
    TSUB(cos2t, 1, tmp0),
    TABS(cos2t, cos2t),
    INSERT_TMACRO(vm::refract, CS(d), CS(eta), CS(ddn), CS(cos2t), CS(tmp0), CS(vtmp))
TKERNEL_END3(d0, d1, d2);

template <std::floating_point T>
TKERNEL2(gtr, dotWmWg, alpha)
    INSERT_TMACRO(vm::gtr, CS(r), CS(dotWmWg), CS(alpha))
TKERNEL_END1(r);

template <std::floating_point T>
TKERNEL2(gaf, ct, alpha)
    INSERT_TMACRO(vm::gaf, CS(r), CS(ct), CS(alpha))
TKERNEL_END1(r);

template <std::floating_point T>
TKERNEL3(slope, ct, u0, u1)
    INSERT_TMACRO(vm::get_slope, CS(r0), CS(r1), CS(ct), CS(u0), CS(u1), CS(tmp0), CS(tmp1), CS(tmp2), CS(tmp3), CS(tmp4), CS(tmp5), CS(tmp6))
TKERNEL_END2(r0, r1);

template <std::floating_point T>
TKERNEL7(microfacet, wi0, wi1, wi2, alpha0, alpha1, u0, u1)
    INSERT_TMACRO(vm::microfacet,
                  CS(r), CS(r0), CS(r1), CS(r2),
                  CS(wi), CS(wi0), CS(wi1), CS(wi2),
                  CS(alpha0), CS(alpha1), CS(u0), CS(u1), CS(s0), CS(s1),
                  CS(tmp0), CS(tmp1), CS(tmp2), CS(tmp3), CS(tmp4), CS(tmp5), CS(tmp6))
TKERNEL_END3(r0, r1, r2);

template <std::floating_point T>
TKERNEL11(rough_i, wi0, wi1, wi2, tu0, tu1, tu2, tv0, tv1, tv2, alpha0, alpha1)
    INSERT_TMACRO(vm::anisorough_i, CS(r), CS(wi), CS(tu), CS(tv), CS(alpha0), CS(alpha1), CS(tmp))
TKERNEL_END1(r);

template <std::floating_point T>
TKERNEL14(rough_n, wm0, wm1, wm2, n0, n1, n2, tu0, tu1, tu2, tv0, tv1, tv2, alpha0, alpha1)
    INSERT_TMACRO(vm::vec3_dot, CS(dotWmWg), CS(wm), CS(n)),
    INSERT_TMACRO(vm::anisorough_n, CS(r), CS(wm), CS(n), CS(tu), CS(tv), CS(dotWmWg), CS(alpha0), CS(alpha1), CS(tmp), CS(vtmp))
TKERNEL_END1(r);

// vector macros of src/vm/lib/linal.hpp
template <std::floating_point T>
TKERNEL6(cross, x0, x1, x2, y0, y1, y2)
    INSERT_TMACRO(vm::vec3_cross, CS(z), CS(x), CS(y), CS(t1))
TKERNEL_END3(z0, z1, z2);

template <std::floating_point T>
TKERNEL12(mat_vec, m00, m01, m02, m10, m11, m12, m20, m21, m22, x0, x1, x2)
    INSERT_TMACRO(vm::mat_vec3, CS(y), CS(m), CS(x))
KERNEL_END(TVAR(y0), TVAR(y1), TVAR(y2));

template <std::floating_point T>
TKERNEL3(norm, v0, v1, v2)
    INSERT_TMACRO(vm::vec3_norm, CS(v), CS(l))
TKERNEL_END3(v0, v1, v2);

// clang-format on
} // namespace loom_ref::pm

/// X(name, kernel type) for every macro kernel (float, as the reference's
/// checks instantiate them)
#define LOOM_REF_CATALOG_MACROS(X)                                      \
    X("pm_disk", loom_ref::pm::disk<float>)                             \
    X("pm_quadratic", loom_ref::pm::quadratic<float>)                   \
    X("pm_sphere", loom_ref::pm::sphere<float>)                         \
    X("pm_cylinder", loom_ref::pm::cylinder<float>)                     \
    X("pm_triangle", loom_ref::pm::triangle<float>)                     \
    X("pm_quad", loom_ref::pm::quad<float>)                             \
    X("pm_onb", loom_ref::pm::onb<float>)                               \
    X("pm_hemisphere", loom_ref::pm::hemisphere<float>)                 \
    X("pm_sample_space", loom_ref::pm::sample_space<float>)             \
    X("pm_ray_dir", loom_ref::pm::ray_dir<float>)                       \
    X("pm_light", loom_ref::pm::light<float>)                           \
    X("pm_schlick", loom_ref::pm::schlick<float>)                       \
    X("pm_reflect", loom_ref::pm::reflect<float>)                       \
    X("pm_refract", loom_ref::pm::refract<float>)                       \
    X("pm_gtr", loom_ref::pm::gtr<float>)                               \
    X("pm_gaf", loom_ref::pm::gaf<float>)                               \
    X("pm_slope", loom_ref::pm::slope<float>)                           \
    X("pm_microfacet", loom_ref::pm::microfacet<float>)                 \
    X("pm_rough_i", loom_ref::pm::rough_i<float>)                       \
    X("pm_rough_n", loom_ref::pm::rough_n<float>)                       \
    X("pm_cross", loom_ref::pm::cross<float>)                           \
    X("pm_mat_vec", loom_ref::pm::mat_vec<float>)                       \
    X("pm_norm", loom_ref::pm::norm<float>)
