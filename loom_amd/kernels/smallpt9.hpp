// -*- C++ -*-
/// @file smallpt9.hpp
/// @brief A smallpt-style path tracer in Loom IR, FP32, one thread per sample.
///
/// Scene: nine spheres — five huge walls/floor/ceiling (radius 1e5; the
/// front is open), a mirror, a glass ball (n = 1.5), a diffuse ball and an
/// emissive light — after Kevin Beason's smallpt scene.  BSDFs: Lambertian
/// (cosine-weighted hemisphere), ideal specular, dielectric with Fresnel
/// (Schlick) and a stochastic reflect/refract choice.  Russian roulette after
/// depth 5, maximum depth 32.  Random numbers: xorshift32 seeded by a hash of
/// the global sample index (pixel·spp + sample).
///
/// Inputs: `sid` (sample index) and `job` = w<<20 | h<<8 | spp.
/// Outputs: the sample radiance (Lx, Ly, Lz); a pixel is the mean of its spp
/// samples in sample order.
///
/// Blocks: START (camera ray), loop (intersect all nine spheres,
/// branch-free), shade (emission, roulette, material), diff, spec, refr.
/// The material blocks jump back to `loop`, so every bounce is a loop trip
/// whose body diverges three ways — the divergence the regrouping backend
/// targets.  Sphere intersection uses the numerically stable quadratic
/// roots q = b + sign(b)·sqrt(det), t = {c/q, q}, which keeps FP32 precise
/// for the 1e5-radius walls.
#pragma once

#include <vm/core.hpp>
#include <vm/macros.hpp>
#include <vm/operators.hpp>

#include <cstdint>

namespace loom_amd::smallpt
{
// clang-format off
// Vector registers are three scalar registers <base>x, <base>y, <base>z.
// The joined names go through one-parameter aliases so that no comma reaches
// the DSL macros.
template <typename V>
using comp_x = loom::intlist::join<V, loom::sl2cs<"x">>;
template <typename V>
using comp_y = loom::intlist::join<V, loom::sl2cs<"y">>;
template <typename V>
using comp_z = loom::intlist::join<V, loom::sl2cs<"z">>;
#define SPT_X(V) comp_x<V>
#define SPT_Y(V) comp_y<V>
#define SPT_Z(V) comp_z<V>

/// s = a · b (fused accumulation)
TMACRO3(vdot, s, a, b)
    TMUL_CS(s, SPT_X(a), SPT_X(b)),
    TFMA_CS(s, SPT_Y(a), SPT_Y(b), s),
    TFMA_CS(s, SPT_Z(a), SPT_Z(b), s)
MACRO_END;

/// v = v / |v| with a reciprocal square root (reference vec3_norm,
/// vm/lib/linal.hpp:453-457)
TMACRO2(vnorm, v, tmp)
    INSERT_TMACRO(vdot, tmp, v, v),
    TRSQRT_CS(tmp, tmp),
    TMUL_CS(SPT_X(v), SPT_X(v), tmp),
    TMUL_CS(SPT_Y(v), SPT_Y(v), tmp),
    TMUL_CS(SPT_Z(v), SPT_Z(v), tmp)
MACRO_END;

/// d = a + b * s (component-wise fused)
TMACRO4(vmadd, d, a, b, s)
    TFMA_CS(SPT_X(d), SPT_X(b), s, SPT_X(a)),
    TFMA_CS(SPT_Y(d), SPT_Y(b), s, SPT_Y(a)),
    TFMA_CS(SPT_Z(d), SPT_Z(b), s, SPT_Z(a))
MACRO_END;

/// xorshift32 step of register rng and u = (rng >> 8) / 2^24 in [0, 1)
TMACRO1(rand01, u)
    ILSHFT(rt, rng, 13),
    IXOR(rng, rng, rt),
    IRSHFT(rt, rng, 17),
    IXOR(rng, rng, rt),
    ILSHFT(rt, rng, 5),
    IXOR(rng, rng, rt),
    IRSHFT(rt, rng, 8),
    MOV(TVAR_CS(u), IVAR(rt)),
    TDIV_CS(u, u, CS(16777216))
MACRO_END;

/// Ray (o, d) against sphere (radius² R2, centre C); keeps the nearest hit
/// distance in tb and the sphere id in bid.
TMACRO5(isect, R2, CX, CY, CZ, ID)
    TSUB_CS(CS(opx), CX, CS(ox)),
    TSUB_CS(CS(opy), CY, CS(oy)),
    TSUB_CS(CS(opz), CZ, CS(oz)),
    INSERT_TMACRO(vdot, CS(b), CS(op), CS(d)),
    INSERT_TMACRO(vdot, CS(c2), CS(op), CS(op)),
    TSUB_CS(CS(c2), CS(c2), R2),
    TFMS(det, b, b, c2),
    TMAX(sq, det, 0),
    TSQRT(sq, sq),
    TCOPYSIGN(sq, sq, b),
    TADD(q, b, sq),
    TDIV(t0, c2, q),
    TMIN(tn, t0, q),
    TMAX(tf, t0, q),
    TGT(okn, tn, 0.01),
    TGT(okf, tf, 0.01),
    TCMOV(t, okn, tn, tf),
    TGE(okd, det, 0),
    ILAND(okf, okf, okd),
    TCMOV(t, okf, t, 1e30),
    TLT(better, t, tb),
    TCMOV(tb, better, t, tb),
    ICMOV_CS(CS(bid), CS(better), ID, CS(bid))
MACRO_END;

// smallpt9: see the file comment.
template <std::floating_point T>
KERNEL(smallpt9, IVAR(sid), IVAR(job))
    // ---- START: decode the sample, seed the generator, camera ray ----
    IRSHFT(w, job, 20),
    IRSHFT(hh, job, 8),
    IBAND(hh, hh, 0xFFF),
    IBAND(spp, job, 0xFF),
    IDIV(pix, sid, spp),
    IMOD(px, pix, w),
    IDIV(py, pix, w),
    IADD(rng, sid, 1),
    IMUL(rng, rng, 0x9E3779B9),
    IRSHFT(rt, rng, 16),
    IXOR(rng, rng, rt),
    IMUL(rng, rng, 0x7FEB352D),
    IRSHFT(rt, rng, 15),
    IXOR(rng, rng, rt),
    IMUL(rng, rng, 0x846CA68B),
    IRSHFT(rt, rng, 16),
    IXOR(rng, rng, rt),
    IEQ(rt, rng, 0),
    ICMOV(rng, rt, 0x9E3779B9, rng),
    MOV(TVAR(fw), IVAR(w)),
    MOV(TVAR(fh), IVAR(hh)),
    INSERT_TMACRO(rand01, CS(u1)),
    INSERT_TMACRO(rand01, CS(u2)),
    MOV(TVAR(sx), IVAR(px)),
    MOV(TVAR(sy), IVAR(py)),
    TADD(sx, sx, u1),
    TADD(sy, sy, u2),
    TDIV(sx, sx, fw),
    TDIV(sy, sy, fh),
    TSUB(sx, sx, 0.5),
    TSUB(sy, sy, 0.5),
    TMUL(ax, fw, 0.5135),
    TDIV(ax, ax, fh),
    TMUL(dx, ax, sx),
    TFMA(dy, sy, 0.5130344317639839, -0.04257336554299295),
    TFMA(dz, sy, -0.021861423206326877, -0.9990933432599491),
    TFMA(ox, dx, 140, 50),
    TFMA(oy, dy, 140, 52),
    TFMA(oz, dz, 140, 295.6),
    INSERT_TMACRO(vnorm, CS(d), CS(len)),
    TMOV(Lx, 0),
    TMOV(Ly, 0),
    TMOV(Lz, 0),
    TMOV(Tx, 1),
    TMOV(Ty, 1),
    TMOV(Tz, 1),
    IMOV(depth, 0),

    // ---- loop: nearest of the nine spheres ----
    LABEL(loop),
    TMOV(tb, 1e30),
    IMOV(bid, 9),
    INSERT_TMACRO(isect, CS(1e10), CS(100001), CS(40.8), CS(81.6), CS(0)),
    INSERT_TMACRO(isect, CS(1e10), CS(-99901), CS(40.8), CS(81.6), CS(1)),
    INSERT_TMACRO(isect, CS(1e10), CS(50), CS(40.8), CS(100000), CS(2)),
    INSERT_TMACRO(isect, CS(1e10), CS(50), CS(100000), CS(81.6), CS(3)),
    INSERT_TMACRO(isect, CS(1e10), CS(50), CS(-99918.4), CS(81.6), CS(4)),
    INSERT_TMACRO(isect, CS(272.25), CS(27), CS(16.5), CS(47), CS(5)),
    INSERT_TMACRO(isect, CS(272.25), CS(73), CS(16.5), CS(78), CS(6)),
    INSERT_TMACRO(isect, CS(72.25), CS(50), CS(8.5), CS(110), CS(7)),
    INSERT_TMACRO(isect, CS(360000), CS(50), CS(681.33), CS(81.6), CS(8)),
    TLT(hit, tb, 1e29),
    IPJMP2(hit, HALT, shade),

    // ---- shade: emission, roulette, material ----
    LABEL(shade),
    INSERT_TMACRO(vmadd, CS(x), CS(o), CS(d), CS(tb)),
    TSCT9(cx, bid, 100001, -99901, 50, 50, 50, 27, 73, 50, 50),
    TSCT9(cy, bid, 40.8, 40.8, 40.8, 100000, -99918.4, 16.5, 16.5, 8.5, 681.33),
    TSCT9(cz, bid, 81.6, 81.6, 100000, 81.6, 81.6, 47, 78, 110, 81.6),
    TSUB(nx, xx, cx),
    TSUB(ny, xy, cy),
    TSUB(nz, xz, cz),
    INSERT_TMACRO(vnorm, CS(n), CS(len)),
    INSERT_TMACRO(vdot, CS(dn), CS(n), CS(d)),
    TLT(into, dn, 0),
    TNEG(mx, nx),
    TNEG(my, ny),
    TNEG(mz, nz),
    TCMOV(nlx, into, nx, mx),
    TCMOV(nly, into, ny, my),
    TCMOV(nlz, into, nz, mz),
    TSCT9(em, bid, 0, 0, 0, 0, 0, 0, 0, 0, 12),
    TFMA(Lx, Tx, em, Lx),
    TFMA(Ly, Ty, em, Ly),
    TFMA(Lz, Tz, em, Lz),
    TSCT9(fx, bid, 0.75, 0.25, 0.75, 0.75, 0.75, 0.999, 0.999, 0.25, 0),
    TSCT9(fy, bid, 0.25, 0.25, 0.75, 0.75, 0.75, 0.999, 0.999, 0.75, 0),
    TSCT9(fz, bid, 0.25, 0.75, 0.75, 0.75, 0.75, 0.999, 0.999, 0.25, 0),
    IADD(depth, depth, 1),
    TMAX(p, fx, fy),
    TMAX(p, p, fz),
    INSERT_TMACRO(rand01, CS(u1)),
    IGT(rr, depth, 5),
    TLT(keep, u1, p),
    ILNOT(alive, rr),
    ILOR(alive, alive, keep),
    ILT(okd, depth, 32),
    ILAND(alive, alive, okd),
    TGT(okp, p, 0),
    ILAND(alive, alive, okp),
    TDIV(ip, 1, p),
    TCMOV(sc, rr, ip, 1),
    TMUL(fx, fx, sc),
    TMUL(fy, fy, sc),
    TMUL(fz, fz, sc),
    TMUL(Tx, Tx, fx),
    TMUL(Ty, Ty, fy),
    TMUL(Tz, Tz, fz),
    ISCT9(mat, bid, 1, 1, 1, 1, 1, 2, 3, 1, 1),
    ICMOV(mat, alive, mat, 0),
    IPJMP4(mat, HALT, diff, spec, refr),

    // ---- diffuse: cosine-weighted hemisphere around nl ----
    LABEL(diff),
    INSERT_TMACRO(rand01, CS(u1)),
    INSERT_TMACRO(rand01, CS(u2)),
    TFMA(r1, u1, 6.2831853, -3.14159265),
    TSQRT(r2s, u2),
    TABS(awx, nlx),
    TGT(big, awx, 0.1),
    TNEG(tmp, nlx),
    TCMOV(ux, big, nlz, 0),
    TNEG(uy, nlz),
    TCMOV(uy, big, 0, uy),
    TCMOV(uz, big, tmp, nly),
    INSERT_TMACRO(vnorm, CS(u), CS(len)),
    TMUL(vx, nlz, uy),
    TFMS(vx, nly, uz, vx),
    TMUL(vy, nlx, uz),
    TFMS(vy, nlz, ux, vy),
    TMUL(vz, nly, ux),
    TFMS(vz, nlx, uy, vz),
    TCOS(cs, r1),
    TSIN(sn, r1),
    TMUL(cs, cs, r2s),
    TMUL(sn, sn, r2s),
    TSUB(wz, 1, u2),
    TSQRT(wz, wz),
    TMUL(dx, ux, cs),
    TMUL(dy, uy, cs),
    TMUL(dz, uz, cs),
    INSERT_TMACRO(vmadd, CS(d), CS(d), CS(v), CS(sn)),
    INSERT_TMACRO(vmadd, CS(d), CS(d), CS(nl), CS(wz)),
    INSERT_TMACRO(vnorm, CS(d), CS(len)),
    TMOV(ox, xx),
    TMOV(oy, xy),
    TMOV(oz, xz),
    JMP(loop),

    // ---- mirror: d - 2 (n·d) n ----
    LABEL(spec),
    TMUL(dn2, dn, -2),
    INSERT_TMACRO(vmadd, CS(d), CS(d), CS(n), CS(dn2)),
    TMOV(ox, xx),
    TMOV(oy, xy),
    TMOV(oz, xz),
    JMP(loop),

    // ---- glass: Fresnel-weighted reflect or refract (total internal
    //      reflection reflects with weight 1) ----
    LABEL(refr),
    TMUL(dn2, dn, -2),
    INSERT_TMACRO(vmadd, CS(r), CS(d), CS(n), CS(dn2)),
    TCMOV(nnt, into, 0.6666666666666666, 1.5),
    INSERT_TMACRO(vdot, CS(ddn), CS(d), CS(nl)),
    TMUL(c2t, ddn, ddn),
    TSUB(c2t, 1, c2t),
    TMUL(nn2, nnt, nnt),
    TMUL(c2t, c2t, nn2),
    TSUB(c2t, 1, c2t),
    TLT(tir, c2t, 0),
    TMAX(sq, c2t, 0),
    TSQRT(sq, sq),
    TFMA(k, ddn, nnt, sq),
    TNEG(km, k),
    TCMOV(k, into, k, km),
    TMUL(tx, dx, nnt),
    TMUL(ty, dy, nnt),
    TMUL(tz, dz, nnt),
    TMUL(ntx, nx, k),
    TMUL(nty, ny, k),
    TMUL(ntz, nz, k),
    TSUB(tx, tx, ntx),
    TSUB(ty, ty, nty),
    TSUB(tz, tz, ntz),
    INSERT_TMACRO(vnorm, CS(t), CS(len)),
    INSERT_TMACRO(vdot, CS(tn), CS(t), CS(n)),
    TNEG(ddm, ddn),
    TCMOV(cc, into, ddm, tn),
    TSUB(cc, 1, cc),
    TMUL(c5, cc, cc),
    TMUL(c5, c5, c5),
    TMUL(c5, c5, cc),
    TFMA(re, c5, 0.96, 0.04),
    TSUB(tr, 1, re),
    TFMA(pp, re, 0.5, 0.25),
    TDIV(rp, re, pp),
    TSUB(tp, 1, pp),
    TDIV(tp, tr, tp),
    INSERT_TMACRO(rand01, CS(u1)),
    TLT(refl, u1, pp),
    ILOR(refl, refl, tir),
    TCMOV(wgt, refl, rp, tp),
    TCMOV(wgt, tir, 1, wgt),
    TMUL(Tx, Tx, wgt),
    TMUL(Ty, Ty, wgt),
    TMUL(Tz, Tz, wgt),
    TCMOV(dx, refl, rx, tx),
    TCMOV(dy, refl, ry, ty),
    TCMOV(dz, refl, rz, tz),
    TMOV(ox, xx),
    TMOV(oy, xy),
    TMOV(oz, xz),
    JMP(loop)
KERNEL_END(TVAR(Lx), TVAR(Ly), TVAR(Lz));

#undef SPT_X
#undef SPT_Y
#undef SPT_Z
// clang-format on
} // namespace loom_amd::smallpt
