// -*- C++ -*-
/// @file ref_macro_checks.cpp
/// @brief ORACLE / TEST INFRASTRUCTURE: the C++ side of the reference's
/// path-tracing and vector macro checks (src/tests/vm_smallpt.cpp:311-700,
/// src/tests/vm_linal.cpp:68-140), for the kernels of ref_macros.hpp.
/// Built by oracle/build_ref.py into oracle/_ref/liboracle_ref.so from the
/// reference checkout.
///
/// Where the reference's C++ compiles here it is called directly: the
/// primitives' intersections, disk sampling and the orthonormal basis
/// (src/smallpt/ref/primitives/*.hpp, transforms.hpp, through niebloids.hpp),
/// the quadratic solver and fma (src/common/math.hpp), vector algebra
/// (src/common/linal.hpp).  The rest of the reference's path-tracer C++
/// lives in src/smallpt/ref/pt.hpp, which cannot be compiled here (it
/// reaches src/common/string.hpp, which needs the absent fast_float
/// library), so the formulas the checks compare with (pt.hpp:48-160,
/// 346-518) are restated below with the reference's own vector types and
/// helpers; each restatement cites the lines it follows.
///
/// ref_pm(name, in, out, n) writes the expected outputs of kernel `name`
/// for n rows of input columns, in the kernel's output layout (an
/// intersection's hit flag as u32); ref_ulp_rows marks the rows whose
/// float outputs all lie within `ulp` of each other by the reference's
/// almost_equal (src/common/math.hpp:253-274).
#include <common/linal.hpp>
#include <common/math.hpp>
#include <smallpt/ref/niebloids.hpp>

#include "ops.hpp" // the op contract's sine and cosine (the reference's SIMD polynomials)

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string_view>

namespace
{
using namespace loom;
using V2 = Vec2<float>;
using V3 = Vec3<float>;

const float*
col(const void* const* in, int k)
{
    return static_cast<const float*>(in[k]);
}

V3
vec3(const void* const* in, int k, std::uint64_t t)
{
    return V3{col(in, k)[t], col(in, k + 1)[t], col(in, k + 2)[t]};
}

void
put3(void* const* out, int k, std::uint64_t t, const V3& v)
{
    for(int i = 0; i < 3; ++i) static_cast<float*>(out[k + i])[t] = v[i];
}

void
put_isect(void* const* out, std::uint64_t t, const Intersection<float>& is)
{
    static_cast<std::uint32_t*>(out[0])[t] = is.intersect ? 1u : 0u;
    static_cast<float*>(out[1])[t]          = is.t;
}

/// A ray as the reference's tests build it: Ray{o, d} normalises d
/// (ray.hpp:36-40); the kernels receive the normalised direction, so it is
/// given here already normalised and kept as is.
Ray<float>
ray_of(const V3& o, const V3& d)
{
    Ray<float> r{o, d};
    r.dir = d;
    return r;
}

// ---- formulas of src/smallpt/ref/pt.hpp, restated ---------------------------
/// pt.hpp:118-128
V3
sample_hemisphere(const V3& n, float u1, float u2)
{
    const float r1    = 2.f * Pi<float> * (u1 - 0.5f);
    const float r2s   = std::sqrt(u2);
    const auto [x, y] = basis(n);
    return normalize(x * std::cos(r1) * r2s + y * std::sin(r1) * r2s + n * std::sqrt(1 - u2));
}

/// pt.hpp:131-153 (ToSampleSpace)
V2
to_sample_space(unsigned x, unsigned y, unsigned w, unsigned h)
{
    const V2    res{static_cast<float>(w), static_cast<float>(h)};
    const float ratio = res[0] / res[1];
    return V2{ratio * (static_cast<float>(x) / res[0] - 0.5f), static_cast<float>(y) / res[1] - 0.5f};
}

/// pt.hpp:155-165
float
schlick(float cd, float eta)
{
    float r0 = (1 - eta) / (1 + eta);
    r0 *= r0;
    return r0 + (1 - r0) * static_pow<5>(1 - cd);
}

/// pt.hpp:42-62 (convert_fov, Perspective), 318-322 (cornell_box::camera)
V3
camera_dir(float ndc0, float ndc1)
{
    const float tan_fov_half = std::tan(radians(110.f) * 0.5f);
    const V3    dir{0.f, 0.f, -1.f};
    return normalize(dir - V3{ndc0 * tan_fov_half, ndc1 * tan_fov_half, 0.f});
}

/// pt.hpp:285-288: the light, a disk of radius 0.35 at (0, 0.9999, 0)
/// facing down; sampled with the reference's disk sampling
/// (primitives/disk.hpp:76-85)
V3
sample_light(float u0, float u1)
{
    Disk<float> light;
    light.position = V3{0.f, 0.9999f, 0.f};
    light.normal   = V3{0.f, -1.f, 0.f};
    light.radius   = 0.35f;
    return sample(V2{u0, u1}, light);
}

/// pt.hpp:348-376: compute_eta, reflect, refract
float
compute_eta(float ior, bool enter)
{
    return enter ? 1 / ior : ior;
}

V3
reflect(const V3& n, const V3& i, float cos_theta)
{
    return i + n * cos_theta * 2.f;
}

V3
refract(const V3& nfn, const V3& i, float eta, float ddn, float cos2t)
{
    return normalize(i * eta - nfn * loom::fma(ddn, eta, std::sqrt(cos2t)));
}

/// pt.hpp:378-401: gtr, gaf
float
gtr(float dotWmWg, float alpha)
{
    return alpha / (alpha + 1 / (dotWmWg * dotWmWg) - 1);
}

float
gaf(float cos_theta, float alpha)
{
    return 2 / (1 + std::sqrt(loom::fma(alpha * alpha, (1 / (cos_theta * cos_theta) - 1), 1.f)));
}

/// Sine and cosine of the slope sampling: the reference's C++ calls
/// std::sin / std::cos (its scalar engine's semantics); the op contract of
/// the VM (and the reference's SIMD engines) evaluates the polynomial of
/// simd.hpp:955-1100, valid on [-pi, pi].  The slope's angle 2*pi*(1 - u)
/// leaves that range, so the two differ there by design; `contract` picks
/// the VM's sine and cosine to check the macro's expansion on its own.
bool g_contract_trig = false;

float
slope_cos(float x)
{
    return g_contract_trig ? loom_oracle::cos_poly(x) : std::cos(x);
}

float
slope_sin(float x)
{
    return g_contract_trig ? loom_oracle::sin_poly(x) : std::sin(x);
}

/// pt.hpp:403-456: slope sampling of the visible normals
V2
get_slope(float cos_theta, float sx, float sy)
{
    if(cos_theta > 0.9999f)
    {
        const float r   = std::sqrt(sx / (1 - sx));
        const float phi = Pi<float> * 2 * (1 - sy);
        return V2{r * slope_cos(phi), r * slope_sin(phi)};
    }
    const float c2    = cos_theta * cos_theta;
    const float B2    = (1 - c2) / std::max(c2, 1e-6f);
    const float B     = std::sqrt(B2);
    const float a     = 1 / B;
    const float G1    = 2 / (1 + std::sqrt(1 + 1 / (a * a)));
    const float A     = 2 * sx / G1 - 1;
    const float A2    = A * A;
    const float den   = A2 - 1;
    const float tmp   = 1 / ((std::abs(den) < 1e-6f) ? std::copysign(1e-6f, den) : den);
    const float D     = std::sqrt(B2 * tmp * tmp - (A2 - B2) * tmp);
    const float Btmp  = B * tmp;
    const float x1    = Btmp - D;
    const float x2    = Btmp + D;
    const float retx  = (A < 0 || x2 > 1 / B) ? x1 : x2;
    float       S     = 1;
    if(sy > .5f)
        sy = 2 * (sy - .5f);
    else
    {
        S  = -1;
        sy = 2 * (.5f - sy);
    }
    const float z = sy * loom::fma(loom::fma(sy, .27385f, -.73369f), sy, .46341f) /
                    loom::fma(loom::fma(loom::fma(sy, .093073f, .309420f), sy, -1.f), sy, .597999f);
    return V2{retx, S * z * std::sqrt(1 + retx * retx)};
}

/// pt.hpp:458-485: microfacet normal of the visible normals
V3
microfacet(V3 wi, float ax, float ay, float u0, float u1)
{
    wi[0] *= ax;
    wi[1] *= ay;
    wi                     = normalize(wi);
    const auto [wx, wy, wz] = wi;
    const V2    s           = get_slope(wz, u0, u1);
    const float lensq       = std::sqrt(wx * wx + wy * wy);
    const bool  tb          = lensq > 1e-12f;
    const float cosphi      = tb ? wx / lensq : 1.f;
    const float sinphi      = tb ? wy / lensq : 0.f;
    const float nx          = loom::fma(sinphi, s[1], cosphi * s[0]) * ax;
    const float ny          = loom::fma(cosphi, s[1], -sinphi * s[0]) * ay;
    return make_normal(-nx, -ny, 1.f);
}

/// pt.hpp:487-518: projected roughness
float
rough_i(const V3& wi, const V3& tu, const V3& tv, float ax, float ay)
{
    const float cs = dot(wi, tu);
    const float sn = dot(wi, tv);
    return std::sqrt(cs * cs * ax * ax + sn * sn * ay * ay);
}

float
rough_n(const V3& wm, const V3& wg, const V3& tu, const V3& tv, float dotWmWg, float ax, float ay)
{
    const V3    ph = normalize(wm - wg * dotWmWg);
    const float cu = dot(ph, tu);
    const float cv = dot(ph, tv);
    return 1 / (cu * cu / ax + cv * cv / ay);
}

template <typename F>
void
rows(std::uint64_t n, F&& f)
{
    for(std::uint64_t t = 0; t < n; ++t) f(t);
}
} // namespace

/// Choose the sine and cosine of the slope sampling: nonzero = the op
/// contract's polynomials, zero = std::sin / std::cos (the default).
extern "C" void
ref_pm_contract_trig(int on)
{
    g_contract_trig = on != 0;
}

extern "C" int
ref_pm(const char* name_c, const void* const* in, void* const* out, std::uint64_t n)
{
    const std::string_view name(name_c);
    auto                   f = [&](int k, std::uint64_t t) { return col(in, k)[t]; };
    auto                   F = [&](int k, std::uint64_t t) -> float&
    { return static_cast<float*>(out[k])[t]; };
    if(name == "pm_disk")
        rows(n, [&](std::uint64_t t) {
            Disk<float> d;
            d.position = vec3(in, 6, t);
            d.normal   = vec3(in, 9, t);
            d.radius   = f(12, t);
            put_isect(out, t, intersect(ray_of(vec3(in, 0, t), vec3(in, 3, t)), d));
        });
    else if(name == "pm_quadratic")
        rows(n, [&](std::uint64_t t) {
            const auto [hit, x] = quadratic(f(0, t), f(1, t), f(2, t));
            static_cast<std::uint32_t*>(out[0])[t] = hit ? 1u : 0u;
            F(1, t)                                 = x;
        });
    else if(name == "pm_sphere")
        rows(n, [&](std::uint64_t t) {
            const Sphere<float> s{vec3(in, 6, t), f(9, t)};
            put_isect(out, t, intersect(ray_of(vec3(in, 0, t), vec3(in, 3, t)), s));
        });
    else if(name == "pm_cylinder")
        rows(n, [&](std::uint64_t t) {
            const Cylinder<float> c{vec3(in, 6, t), f(9, t), f(10, t)};
            put_isect(out, t, intersect(ray_of(vec3(in, 0, t), vec3(in, 3, t)), c));
        });
    else if(name == "pm_triangle")
        rows(n, [&](std::uint64_t t) {
            const Triangle<float> tri{vec3(in, 6, t), vec3(in, 9, t), vec3(in, 12, t)};
            put_isect(out, t, intersect(ray_of(vec3(in, 0, t), vec3(in, 3, t)), tri));
        });
    else if(name == "pm_quad")
        rows(n, [&](std::uint64_t t) {
            const Quad<float> q{vec3(in, 6, t), vec3(in, 9, t), vec3(in, 12, t), vec3(in, 15, t)};
            put_isect(out, t, intersect(ray_of(vec3(in, 0, t), vec3(in, 3, t)), q));
        });
    else if(name == "pm_onb")
        rows(n, [&](std::uint64_t t) {
            const auto [b, tt] = basis(vec3(in, 0, t));
            put3(out, 0, t, b);
            put3(out, 3, t, tt);
        });
    else if(name == "pm_hemisphere")
        rows(n, [&](std::uint64_t t) { put3(out, 0, t, sample_hemisphere(vec3(in, 0, t), f(3, t), f(4, t))); });
    else if(name == "pm_sample_space")
        rows(n, [&](std::uint64_t t) {
            auto u        = [&](int k) { return static_cast<const std::uint32_t*>(in[k])[t]; };
            const V2 s    = to_sample_space(u(0), u(1), u(2), u(3));
            F(0, t)       = s[0];
            F(1, t)       = s[1];
        });
    else if(name == "pm_ray_dir")
        rows(n, [&](std::uint64_t t) { put3(out, 0, t, camera_dir(f(0, t), f(1, t))); });
    else if(name == "pm_light")
        rows(n, [&](std::uint64_t t) { put3(out, 0, t, sample_light(f(0, t), f(1, t))); });
    else if(name == "pm_schlick")
        rows(n, [&](std::uint64_t t) { F(0, t) = schlick(f(0, t), f(1, t)); });
    else if(name == "pm_reflect")
        rows(n, [&](std::uint64_t t) {
            const V3 nn = vec3(in, 0, t), i = vec3(in, 3, t);
            put3(out, 0, t, reflect(nn, i, dot(nn, -i)));
        });
    else if(name == "pm_refract")
        // the reference test's own composition (vm_smallpt.cpp:556-567)
        rows(n, [&](std::uint64_t t) {
            const V3    nn        = vec3(in, 0, t), i = vec3(in, 3, t);
            const float cos_theta = dot(nn, -i);
            const bool  into      = cos_theta > 0;
            const V3    nfn       = into ? nn : -nn;
            const float eta       = compute_eta(1.5f, into);
            const float ddn       = dot(i, nfn);
            const float cos2t     = 1 - eta * eta * (1 - ddn * ddn);
            put3(out, 0, t, refract(nfn, i, eta, ddn, std::abs(cos2t)));
        });
    else if(name == "pm_gtr")
        rows(n, [&](std::uint64_t t) { F(0, t) = gtr(f(0, t), f(1, t)); });
    else if(name == "pm_gaf")
        rows(n, [&](std::uint64_t t) { F(0, t) = gaf(f(0, t), f(1, t)); });
    else if(name == "pm_slope")
        rows(n, [&](std::uint64_t t) {
            const V2 s = get_slope(f(0, t), f(1, t), f(2, t));
            F(0, t)    = s[0];
            F(1, t)    = s[1];
        });
    else if(name == "pm_microfacet")
        rows(n, [&](std::uint64_t t) {
            put3(out, 0, t, microfacet(vec3(in, 0, t), f(3, t), f(4, t), f(5, t), f(6, t)));
        });
    else if(name == "pm_rough_i")
        rows(n, [&](std::uint64_t t) {
            F(0, t) = rough_i(vec3(in, 0, t), vec3(in, 3, t), vec3(in, 6, t), f(9, t), f(10, t));
        });
    else if(name == "pm_rough_n")
        rows(n, [&](std::uint64_t t) {
            const V3 wm = vec3(in, 0, t), wg = vec3(in, 3, t);
            F(0, t) = rough_n(wm, wg, vec3(in, 6, t), vec3(in, 9, t), dot(wm, wg), f(12, t), f(13, t));
        });
    else if(name == "pm_cross")
        rows(n, [&](std::uint64_t t) { put3(out, 0, t, cross(vec3(in, 0, t), vec3(in, 3, t))); });
    else if(name == "pm_mat_vec")
        rows(n, [&](std::uint64_t t) {
            std::array<float, 9> rowmajor{};
            for(int k = 0; k < 9; ++k) rowmajor[k] = f(k, t);
            const Mat3<float> m{rowmajor};
            put3(out, 0, t, dot(m, vec3(in, 9, t)));
        });
    else if(name == "pm_norm")
        rows(n, [&](std::uint64_t t) { put3(out, 0, t, normalize(vec3(in, 0, t))); });
    else
        return 1;
    return 0;
}

/// ok[t] = 1 when every one of the k float columns a[j][t], b[j][t] agree
/// within ulp by the reference's almost_equal.
extern "C" void
ref_ulp_rows(const float* const* a, const float* const* b, int k, std::uint64_t n, unsigned ulp,
             std::uint8_t* ok)
{
    for(std::uint64_t t = 0; t < n; ++t)
    {
        bool all = true;
        for(int j = 0; j < k; ++j) all = all && loom::almost_equal(a[j][t], b[j][t], ulp);
        ok[t] = all ? 1 : 0;
    }
}
