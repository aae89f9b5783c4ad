// -*- C++ -*-
/// @file smallpt9_direct.hpp
/// @brief ORACLE (test infrastructure only): the 9-sphere path tracer written
/// directly in C++, operation for operation as the IR kernel specifies it
/// (loom_amd/kernels/smallpt9.hpp), with the op contract of ops.hpp
/// (no contraction, fused fma where the IR says fma, polynomial sin/cos,
/// 1/sqrt for rsqrt, literal values from the reference literal parser).
///
/// It checks the IR kernel and the interpreter from outside the IR, and it
/// is the compiled scalar CPU baseline the GPU is timed against.
#pragma once

#include "ops.hpp"

#include <cmath>
#include <cstdint>

namespace loom_oracle::direct
{
/// Literal values exactly as the IR front-end reads them.
struct Lits
{
    float v(const char* s) const { return *lit::parse<float>(s); }
};

struct Sphere
{
    float r2, cx, cy, cz;
};

struct Scene
{
    Sphere s[9];
    float  cx[9], cy[9], cz[9], em[9], fx[9], fy[9], fz[9];
    std::uint32_t mat[9];
    float  eps, inf, hit_max, half, a5135, cyy, camdy, cyz, camdz, k140, o50, o52, o2956;
    float  two_pi, m_pi, tenth, nnt_in, nnt_out, f096, f004, half_f, quarter, m2, i2p24;
    float  zero, one;

    Scene()
    {
        auto P = [](const char* t) { return *lit::parse<float>(t); };
        const char* r2[9]  = {"1e10", "1e10", "1e10", "1e10", "1e10", "272.25", "272.25", "72.25", "360000"};
        const char* sx[9]  = {"100001", "-99901", "50", "50", "50", "27", "73", "50", "50"};
        const char* sy[9]  = {"40.8", "40.8", "40.8", "100000", "-99918.4", "16.5", "16.5", "8.5", "681.33"};
        const char* sz[9]  = {"81.6", "81.6", "100000", "81.6", "81.6", "47", "78", "110", "81.6"};
        const char* ems[9] = {"0", "0", "0", "0", "0", "0", "0", "0", "12"};
        const char* fxs[9] = {"0.75", "0.25", "0.75", "0.75", "0.75", "0.999", "0.999", "0.25", "0"};
        const char* fys[9] = {"0.25", "0.25", "0.75", "0.75", "0.75", "0.999", "0.999", "0.75", "0"};
        const char* fzs[9] = {"0.25", "0.75", "0.75", "0.75", "0.75", "0.999", "0.999", "0.25", "0"};
        const std::uint32_t mats[9] = {1, 1, 1, 1, 1, 2, 3, 1, 1};
        for(int i = 0; i < 9; ++i)
        {
            s[i]   = {P(r2[i]), P(sx[i]), P(sy[i]), P(sz[i])};
            cx[i]  = P(sx[i]);
            cy[i]  = P(sy[i]);
            cz[i]  = P(sz[i]);
            em[i]  = P(ems[i]);
            fx[i]  = P(fxs[i]);
            fy[i]  = P(fys[i]);
            fz[i]  = P(fzs[i]);
            mat[i] = mats[i];
        }
        eps = P("0.01"); inf = P("1e30"); hit_max = P("1e29"); half = P("0.5");
        a5135 = P("0.5135"); cyy = P("0.5130344317639839"); camdy = P("-0.04257336554299295");
        cyz = P("-0.021861423206326877"); camdz = P("-0.9990933432599491");
        k140 = P("140"); o50 = P("50"); o52 = P("52"); o2956 = P("295.6");
        two_pi = P("6.2831853"); m_pi = P("-3.14159265"); tenth = P("0.1");
        nnt_in = P("0.6666666666666666"); nnt_out = P("1.5"); f096 = P("0.96"); f004 = P("0.04");
        half_f = P("0.5"); quarter = P("0.25"); m2 = P("-2"); i2p24 = P("16777216");
        zero = 0.0f; one = 1.0f;
    }
};

inline float
fmaxs(float a, float b) // std::max argument order (operators.hpp:314)
{
    return (a < b) ? b : a;
}

inline float
fmins(float a, float b)
{
    return (b < a) ? b : a;
}

struct Rng
{
    std::uint32_t s;
    float next(const Scene& S)
    {
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        return static_cast<float>(s >> 8) / S.i2p24;
    }
};

inline float
dot(float ax, float ay, float az, float bx, float by, float bz)
{
    float s = ax * bx;
    s       = std::fma(ay, by, s);
    return std::fma(az, bz, s);
}

inline void
norm(float& x, float& y, float& z)
{
    const float l = 1.0f / std::sqrt(dot(x, y, z, x, y, z));
    x *= l;
    y *= l;
    z *= l;
}

/// Radiance of sample @p sid of frame word @p job (w<<20 | h<<8 | spp).
inline void
sample(const Scene& S, std::uint32_t sid, std::uint32_t job, float& Lx, float& Ly, float& Lz)
{
    const std::uint32_t w = job >> 20, hh = (job >> 8) & 0xFFFu, spp = job & 0xFFu;
    const std::uint32_t pix = spp ? sid / spp : ~0u;
    const std::uint32_t px = w ? pix % w : pix, py = w ? pix / w : ~0u;
    std::uint32_t       r  = (sid + 1) * 0x9E3779B9u;
    r ^= r >> 16;
    r *= 0x7FEB352Du;
    r ^= r >> 15;
    r *= 0x846CA68Bu;
    r ^= r >> 16;
    if(r == 0) r = 0x9E3779B9u;
    Rng         R{r};
    const float fw = static_cast<float>(w), fh = static_cast<float>(hh);
    const float u1 = R.next(S), u2 = R.next(S);
    float       sx = static_cast<float>(px), sy = static_cast<float>(py);
    sx = sx + u1;
    sy = sy + u2;
    sx = sx / fw;
    sy = sy / fh;
    sx = sx - S.half;
    sy = sy - S.half;
    float ax = fw * S.a5135;
    ax       = ax / fh;
    float dx = ax * sx;
    float dy = std::fma(sy, S.cyy, S.camdy);
    float dz = std::fma(sy, S.cyz, S.camdz);
    float ox = std::fma(dx, S.k140, S.o50);
    float oy = std::fma(dy, S.k140, S.o52);
    float oz = std::fma(dz, S.k140, S.o2956);
    norm(dx, dy, dz);
    Lx = Ly = Lz = 0.0f;
    float         Tx = 1.0f, Ty = 1.0f, Tz = 1.0f;
    std::uint32_t depth = 0;
    for(;;)
    {
        // loop: nearest sphere
        float         tb  = S.inf;
        std::uint32_t bid = 9;
        for(int i = 0; i < 9; ++i)
        {
            const float opx = S.s[i].cx - ox, opy = S.s[i].cy - oy, opz = S.s[i].cz - oz;
            const float b   = dot(opx, opy, opz, dx, dy, dz);
            float       c2  = dot(opx, opy, opz, opx, opy, opz);
            c2              = c2 - S.s[i].r2;
            const float det = std::fma(b, b, -c2);
            float       sq  = fmaxs(det, 0.0f);
            sq              = std::sqrt(sq);
            sq              = std::copysign(sq, b);
            const float q   = b + sq;
            const float t0  = c2 / q;
            const float tn = fmins(t0, q), tf = fmaxs(t0, q);
            const bool  okn = tn > S.eps;
            bool        okf = tf > S.eps;
            float       t   = okn ? tn : tf;
            okf             = okf && (det >= 0.0f);
            t               = okf ? t : S.inf;
            const bool better = t < tb;
            tb                = better ? t : tb;
            bid               = better ? static_cast<std::uint32_t>(i) : bid;
        }
        if(!(tb < S.hit_max)) return;
        // shade
        const float xx = std::fma(dx, tb, ox), xy = std::fma(dy, tb, oy), xz = std::fma(dz, tb, oz);
        auto sel = [&](const float* a) { return bid < 9 ? a[bid] : 0.0f; };
        float nx = xx - sel(S.cx), ny = xy - sel(S.cy), nz = xz - sel(S.cz);
        norm(nx, ny, nz);
        const float dn   = dot(nx, ny, nz, dx, dy, dz);
        const bool  into = dn < 0.0f;
        const float nlx = into ? nx : -nx, nly = into ? ny : -ny, nlz = into ? nz : -nz;
        const float em = sel(S.em);
        Lx             = std::fma(Tx, em, Lx);
        Ly             = std::fma(Ty, em, Ly);
        Lz             = std::fma(Tz, em, Lz);
        float fx = sel(S.fx), fy = sel(S.fy), fz = sel(S.fz);
        depth += 1;
        float p = fmaxs(fx, fy);
        p       = fmaxs(p, fz);
        const float u  = R.next(S);
        const bool  rr = depth > 5;
        bool alive = (!rr || u < p) && depth < 32 && p > 0.0f;
        const float ip = 1.0f / p;
        const float sc = rr ? ip : 1.0f;
        fx *= sc;
        fy *= sc;
        fz *= sc;
        Tx *= fx;
        Ty *= fy;
        Tz *= fz;
        std::uint32_t mat = bid < 9 ? S.mat[bid] : 0u;
        if(!alive) mat = 0;
        if(mat == 0) return;
        if(mat == 1)
        {
            const float v1 = R.next(S), v2 = R.next(S);
            const float r1  = std::fma(v1, S.two_pi, S.m_pi);
            const float r2s = std::sqrt(v2);
            const bool  big = std::fabs(nlx) > S.tenth;
            float ux = big ? nlz : 0.0f;
            float uy = -nlz;
            uy       = big ? 0.0f : uy;
            float uz = big ? -nlx : nly;
            norm(ux, uy, uz);
            float vx = nlz * uy;
            vx       = std::fma(nly, uz, -vx);
            float vy = nlx * uz;
            vy       = std::fma(nlz, ux, -vy);
            float vz = nly * ux;
            vz       = std::fma(nlx, uy, -vz);
            float cs = cos_poly(r1), sn = sin_poly(r1);
            cs       = cs * r2s;
            sn       = sn * r2s;
            float wz = 1.0f - v2;
            wz       = std::sqrt(wz);
            dx       = ux * cs;
            dy       = uy * cs;
            dz       = uz * cs;
            dx       = std::fma(vx, sn, dx);
            dy       = std::fma(vy, sn, dy);
            dz       = std::fma(vz, sn, dz);
            dx       = std::fma(nlx, wz, dx);
            dy       = std::fma(nly, wz, dy);
            dz       = std::fma(nlz, wz, dz);
            norm(dx, dy, dz);
        }
        else if(mat == 2)
        {
            const float dn2 = dn * S.m2;
            dx              = std::fma(nx, dn2, dx);
            dy              = std::fma(ny, dn2, dy);
            dz              = std::fma(nz, dn2, dz);
        }
        else
        {
            const float dn2 = dn * S.m2;
            const float rx = std::fma(nx, dn2, dx), ry = std::fma(ny, dn2, dy),
                        rz = std::fma(nz, dn2, dz);
            const float nnt = into ? S.nnt_in : S.nnt_out;
            const float ddn = dot(dx, dy, dz, nlx, nly, nlz);
            float       c2t = ddn * ddn;
            c2t             = 1.0f - c2t;
            const float nn2 = nnt * nnt;
            c2t             = c2t * nn2;
            c2t             = 1.0f - c2t;
            const bool tir  = c2t < 0.0f;
            float      sq   = fmaxs(c2t, 0.0f);
            sq              = std::sqrt(sq);
            float k         = std::fma(ddn, nnt, sq);
            k               = into ? k : -k;
            float tx = dx * nnt, ty = dy * nnt, tz = dz * nnt;
            tx -= nx * k;
            ty -= ny * k;
            tz -= nz * k;
            norm(tx, ty, tz);
            const float tn = dot(tx, ty, tz, nx, ny, nz);
            float       cc = into ? -ddn : tn;
            cc             = 1.0f - cc;
            float c5       = cc * cc;
            c5             = c5 * c5;
            c5             = c5 * cc;
            const float re = std::fma(c5, S.f096, S.f004);
            const float tr = 1.0f - re;
            const float pp = std::fma(re, S.half_f, S.quarter);
            const float rp = re / pp;
            float       tp = 1.0f - pp;
            tp             = tr / tp;
            const float u3 = R.next(S);
            bool        refl = u3 < pp;
            refl             = refl || tir;
            float wgt        = refl ? rp : tp;
            wgt              = tir ? 1.0f : wgt;
            Tx *= wgt;
            Ty *= wgt;
            Tz *= wgt;
            dx = refl ? rx : tx;
            dy = refl ? ry : ty;
            dz = refl ? rz : tz;
        }
        ox = xx;
        oy = xy;
        oz = xz;
    }
}
} // namespace loom_oracle::direct
