// -*- C++ -*-
/// @file programs.hpp
/// @brief Integer and small floating-point kernels in Loom IR.
///
/// `fib` is the README kernel of the reference (README.md:133-151, the same
/// program as tests/programs.hpp:201-218): it must compile unchanged, so its
/// statements are the DSL text of that listing.  The other kernels exercise
/// each part of the op contract and of the block structure; their expected
/// results are closed forms checked by the tests (reference validators:
/// tests/programs.hpp:288-389).
#pragma once

#include <vm/core.hpp>
#include <vm/macros.hpp>
#include <vm/operators.hpp>

#include <cstdint>

namespace loom_amd::programs
{
// clang-format off

// Fibonacci number F(n) (README.md:133-151).
template <std::unsigned_integral T>
TKERNEL1(fib, n)
    TLE(cmp, n, 1),
    TCMOV(c, cmp, n, 1),
    IPJMP2(cmp, work, HALT),
    LABEL(work),
    TMOV(p, 0),
    TMOV(c, 1),
    TMOV(i, 2),
    LABEL(loop),
    TADD(t, p, c),
    TMOV(p, c),
    TMOV(c, t),
    TADD(i, i, 1),
    TLE(cmp, i, n),
    IPJMP2(cmp, HALT, loop)
TKERNEL_END1(c);

// Collatz total stopping time: number of x -> x/2 (even) or 3x+1 (odd) steps
// until x == 1; 0 for x <= 1.  A loop whose body branches on parity, so lanes
// of a wave diverge on every iteration and finish after very different trip
// counts.  The count leaves the kernel as a u32 register (a second register
// width next to the u64 state).
template <std::unsigned_integral T>
TKERNEL1(collatz, x)
    IMOV(cnt, 0),
    ILE(fin, x, 1),
    IPJMP2(fin, step, done),
    LABEL(step),
    IBAND(odd, x, 1),
    IPJMP2(odd, even, odd3),
    LABEL(even),
    IRSHFT(x, x, 1),
    JMP(next),
    LABEL(odd3),
    IMUL(x, x, 3),
    IADD(x, x, 1),
    LABEL(next),
    IADD(cnt, cnt, 1),
    IEQ(fin, x, 1),
    IPJMP2(fin, step, done),
    LABEL(done),
    MOV(VAR(std::uint32_t, steps), IVAR(cnt))
KERNEL_END(VAR(std::uint32_t, steps));

// a! for a <= 20 (u64), 1 for a <= 1: a loop counting down.
template <std::unsigned_integral T>
TKERNEL1(factorial, a)
    TLE(small, a, 1),
    TCMOV(r, small, 1, a),
    IPJMP2(small, init, HALT),
    LABEL(init),
    TSUB(k, a, 1),
    LABEL(body),
    TMUL(r, r, k),
    TSUB(k, k, 1),
    TEQ(z, k, 0),
    IPJMP2(z, body, HALT)
TKERNEL_END1(r);

// Four-way jump table: i in {0,1,2,3} -> {1,3,7,11}; other values halt at
// once (out-of-range ParmJump index -> HALT) and return i unchanged.
template <typename T>
KERNEL(jump4, IVAR(i))
    IPJMP4(i, c0, c1, c2, c3),
    LABEL(c0),
    IMOV(i, 1),
    JMP(HALT),
    LABEL(c1),
    IMOV(i, 3),
    JMP(HALT),
    LABEL(c2),
    IMOV(i, 7),
    JMP(HALT),
    LABEL(c3),
    IMOV(i, 11)
KERNEL_END(IVAR(i));

// Integer op mix, straight-line: every binary integer operation of the
// contract on (a, b), folded into several outputs.
template <std::unsigned_integral T>
TKERNEL2(int_mix, a, b)
    TADD(s, a, b),
    TSUB(d, a, b),
    TMUL(m, a, b),
    TDIV(q, a, b),
    TMOD(r, a, b),
    IBAND(x1, a, b),
    IBOR(x2, a, b),
    IXOR(x3, a, b),
    IBNOT(x4, a),
    ILSHFT(l, a, b),
    IRSHFT(h, a, b),
    IMIN(mn, a, b),
    IMAX(mx, a, b),
    TNEG(ng, a),
    ILAND(la, a, b),
    ILOR(lo, a, b),
    ILNOT(ln, a),
    TLT(c1, a, b),
    TGE(c2, a, b),
    TNEQ(c3, a, b),
    ISCT4(sc, c1, 10, 0x20, 0b11, 4),
    IANDLT(p1, c2, a, b),
    IORGE(p2, c1, a, b),
    IXOR(x1, x1, x2),
    IXOR(x1, x1, x3),
    IXOR(x1, x1, x4),
    IADD(s, s, l),
    IADD(s, s, h),
    IADD(d, d, mn),
    IADD(d, d, mx),
    IADD(m, m, ng),
    IADD(q, q, r),
    IADD(la, la, lo),
    IADD(la, la, ln),
    IADD(la, la, c1),
    IADD(la, la, c2),
    IADD(la, la, c3),
    IADD(la, la, sc),
    IADD(la, la, p1),
    IADD(la, la, p2)
TKERNEL_END6(s, d, m, q, x1, la);

// Sine and cosine of x: the op contract's polynomials (the reference SIMD
// backends', simd.hpp:955-1100), valid on [-pi, pi].  Checked bit for bit
// against vectors evaluated in exact arithmetic from the reference's tables.
template <std::floating_point T>
TKERNEL1(sincos, x)
    TSIN(s, x),
    TCOS(c, x)
TKERNEL_END2(s, c);

// Square root and reciprocal square root, one of each per input: the
// correctly rounded results of the op contract over the whole input range
// (zeros, subnormals, infinities, NaN and negatives included).
template <std::floating_point T>
TKERNEL1(roots, x)
    TSQRT(s, x),
    TRSQRT(r, x)
TKERNEL_END2(s, r);

// Division by literals: powers of two (whose quotients the backend forms as
// products with the exact reciprocal) beside divisors that are not.
template <std::floating_point T>
TKERNEL1(literal_div, x)
    TDIV(q1, x, 0.5),
    TDIV(q2, x, 16777216),
    TDIV(q3, x, -0.125),
    TDIV(q4, x, 3),
    TDIV(q5, x, 0.1),
    TDIV(q6, 1, x)
TKERNEL_END6(q1, q2, q3, q4, q5, q6);

// Constant selection with indices inside and outside the constant list:
// sct(i, c0, ..., cN-1) is c_i for i < N and 0 otherwise (operators.hpp
// sct).  The backend reads a list of at least four 4-byte constants from a
// table in LDS and the others through compares and selects; this kernel has
// lists of both kinds.
template <std::floating_point T>
KERNEL(sct_range, IVAR(x))
    TSCT6(a, x, 1.5, -2.25, 3e3, 0.125, -7, 9.5),
    ISCT5(b, x, 7, 0xFFFF, 3, 0x12345678, 1),
    ISUB(y, x, 2),
    TSCT4(c, y, 0.5, 0.25, -1, 2),
    TSCT3(s3, x, 4, 5, 6)
KERNEL_END(TVAR(a), IVAR(b), TVAR(c), TVAR(s3));

// Float op mix, straight-line, exercising the floating-point contract:
// arithmetic, fused multiply-add/sub, sqrt/rsqrt, rounding, min/max,
// copysign, sin/cos polynomials, select, union register and conversions.
template <std::floating_point T>
TKERNEL2(fp_mix, x, y)
    TMUL(a, x, y),
    TADD(a, a, x),
    TDIV(b, x, y),
    TFMA(c, x, y, b),
    TFMS(d, x, y, a),
    TABS(g1, x),
    TSQRT(g1, g1),
    TRSQRT(g2, g1),
    TROUND(r1, b),
    TFLOOR(r2, b),
    TCEIL(r3, b),
    TTRUNC(r4, b),
    TADD(r1, r1, r2),
    TADD(r3, r3, r4),
    TMIN(mn, x, y),
    TMAX(mx, x, y),
    TCOPYSIGN(cs, x, y),
    TMUL(ang, x, 0.001),
    TSIN(sn, ang),
    TCOS(cn, ang),
    TLT(sel, x, y),
    TCMOV(pick, sel, mn, mx),
    TSCT3(k, sel, 0.5, 1.25, -3e2),
    TADD(a, a, c),
    TADD(a, a, d),
    TADD(b, b, g2),
    TADD(b, b, r1),
    TADD(b, b, r3),
    TADD(cs, cs, pick),
    TADD(cs, cs, k),
    TMUL(sn, sn, cn),
    MOV(IVAR(u), TVAR(x)),
    TMOV(sum, a),
    TADD(sum, sum, b)
KERNEL_END(TVAR(a), TVAR(b), TVAR(cs), TVAR(sn), IVAR(u), TVAR(sum));

// clang-format on
} // namespace loom_amd::programs
