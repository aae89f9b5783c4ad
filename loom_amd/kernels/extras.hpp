// -*- C++ -*-
/// @file extras.hpp
/// @brief Two more workloads of the kind the reference benchmarks: a
/// straight-line counter-based generator and a divergent floating-point
/// root finder.
///
/// * `threefry2x32` — Threefry-2x32 with 32 rounds (Salmon, Moraes, Dror and
///   Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011): key
///   (0xdead, 0xbeef), rotation schedule 13 15 26 6 / 17 29 16 24, a key
///   injection after every four rounds.  The reference benchmarks the same
///   generator with the same key (benchmarks/threefry.cpp:33-43, macros
///   vm/lib/random.hpp:30-75).  Outputs the two words of the block and the
///   counter advanced by 39916801 as a 64-bit number (c1:c0).
/// * `cubic_bisect` — a root of a·x³ + b·x² + c·x + d by bisection, the
///   algorithm of the reference's root-finding test (tests/root_finding.hpp:
///   31-119, validator 193-216): bracket the root by widening [-10, 10]
///   tenfold until the signs at the ends differ, then halve until
///   |f(mid)| < 0.1.  Outputs the midpoint and the number of halvings.  Unlike
///   the reference, both loops are capped (40 widenings, 256 halvings): an
///   input whose root cannot be met to 0.1 in floating point would otherwise
///   spin forever, and a GPU kernel must end.
#pragma once

#include <vm/core.hpp>
#include <vm/macros.hpp>
#include <vm/operators.hpp>

#include <cstdint>

namespace loom_amd::programs
{
// clang-format off

/// One round: x0 += x1; x1 = rotl(x1, L) ^ x0, with R = 32 - L.
TMACRO2(tf_round, L, R)
    IADD_CS(CS(x0), CS(x0), CS(x1)),
    ILSHFT_CS(CS(rl), CS(x1), L),
    IRSHFT_CS(CS(rr), CS(x1), R),
    IBOR_CS(CS(x1), CS(rl), CS(rr)),
    IXOR_CS(CS(x1), CS(x1), CS(x0))
MACRO_END;

/// Four rounds with rotations L1..L4 (Ri = 32 - Li).
TMACRO8(tf_rounds4, L1, R1, L2, R2, L3, R3, L4, R4)
    INSERT_TMACRO(tf_round, L1, R1),
    INSERT_TMACRO(tf_round, L2, R2),
    INSERT_TMACRO(tf_round, L3, R3),
    INSERT_TMACRO(tf_round, L4, R4)
MACRO_END;

/// Key injection number S: x0 += K0; x1 += K1 + S.
TMACRO3(tf_inject, K0, K1, S)
    IADD_CS(CS(x0), CS(x0), K0),
    IADD_CS(CS(x1), CS(x1), K1),
    IADD_CS(CS(x1), CS(x1), S)
MACRO_END;

// key words: k0 = 0xdead, k1 = 0xbeef, k2 = 0x1bd11bda ^ k0 ^ k1 = 0x1bd17b98
template <std::unsigned_integral T>
TKERNEL2(threefry2x32, c0, c1)
    IADD(x0, c0, 0xdead),
    IADD(x1, c1, 0xbeef),
    INSERT_TMACRO(tf_rounds4, CS(13), CS(19), CS(15), CS(17), CS(26), CS(6), CS(6), CS(26)),
    INSERT_TMACRO(tf_inject, CS(0xbeef), CS(0x1bd17b98), CS(1)),
    INSERT_TMACRO(tf_rounds4, CS(17), CS(15), CS(29), CS(3), CS(16), CS(16), CS(24), CS(8)),
    INSERT_TMACRO(tf_inject, CS(0x1bd17b98), CS(0xdead), CS(2)),
    INSERT_TMACRO(tf_rounds4, CS(13), CS(19), CS(15), CS(17), CS(26), CS(6), CS(6), CS(26)),
    INSERT_TMACRO(tf_inject, CS(0xdead), CS(0xbeef), CS(3)),
    INSERT_TMACRO(tf_rounds4, CS(17), CS(15), CS(29), CS(3), CS(16), CS(16), CS(24), CS(8)),
    INSERT_TMACRO(tf_inject, CS(0xbeef), CS(0x1bd17b98), CS(4)),
    INSERT_TMACRO(tf_rounds4, CS(13), CS(19), CS(15), CS(17), CS(26), CS(6), CS(6), CS(26)),
    INSERT_TMACRO(tf_inject, CS(0x1bd17b98), CS(0xdead), CS(5)),
    INSERT_TMACRO(tf_rounds4, CS(17), CS(15), CS(29), CS(3), CS(16), CS(16), CS(24), CS(8)),
    INSERT_TMACRO(tf_inject, CS(0xdead), CS(0xbeef), CS(6)),
    INSERT_TMACRO(tf_rounds4, CS(13), CS(19), CS(15), CS(17), CS(26), CS(6), CS(6), CS(26)),
    INSERT_TMACRO(tf_inject, CS(0xbeef), CS(0x1bd17b98), CS(7)),
    INSERT_TMACRO(tf_rounds4, CS(17), CS(15), CS(29), CS(3), CS(16), CS(16), CS(24), CS(8)),
    INSERT_TMACRO(tf_inject, CS(0x1bd17b98), CS(0xdead), CS(8)),
    // the counter as a 64-bit number c1:c0, advanced by 39916801
    IADD(n0, c0, 39916801),
    ILT(carry, n0, c0),
    IADD(n1, c1, carry)
TKERNEL_END4(x0, x1, n0, n1);

/// Y = a·X³ + b·X² + c·X + d as fma(a, X³, fma(b, X², fma(c, X, d))).
TMACRO2(cubic_at, Y, X)
    TMUL_CS(CS(x2), X, X),
    TMUL_CS(CS(x3), CS(x2), X),
    TFMA_CS(Y, CS(c), X, CS(d)),
    TFMA_CS(Y, CS(b), CS(x2), Y),
    TFMA_CS(Y, CS(a), CS(x3), Y)
MACRO_END;

/// F = 1 when the cubic is negative at exactly one of P and Q.
TMACRO3(sign_change, F, P, Q)
    INSERT_TMACRO(cubic_at, CS(fp), P),
    INSERT_TMACRO(cubic_at, CS(fq), Q),
    TLT_CS(CS(fp), CS(fp), CS(0)),
    TLT_CS(CS(fq), CS(fq), CS(0)),
    INEQ_CS(F, CS(fp), CS(fq))
MACRO_END;

/// M = lo + (hi - lo) · 0.5
TMACRO1(midpoint, M)
    TSUB_CS(M, CS(hi), CS(lo)),
    TMUL_CS(M, M, CS(0.5)),
    TADD_CS(M, M, CS(lo))
MACRO_END;

template <std::floating_point T>
TKERNEL4(cubic_bisect, a, b, c, d)
    TMOV(lo, -10),
    TMOV(hi, 10),
    IMOV(w, 0),
    LABEL(widen),
    INSERT_TMACRO(sign_change, CS(split), CS(lo), CS(hi)),
    IADD(w, w, 1),
    ILT(wmore, w, 40),
    INEQ(stay, split, 1),
    ILAND(stay, stay, wmore),
    IPJMP2(stay, bisect, grow),
    LABEL(grow),
    TMUL(lo, lo, 10),
    TMUL(hi, hi, 10),
    JMP(widen),
    LABEL(bisect),
    INSERT_TMACRO(midpoint, CS(mid)),
    INSERT_TMACRO(cubic_at, CS(fm), CS(mid)),
    IMOV(i, 0),
    TABS(fm, fm),
    TGE(far, fm, 0.1),
    IPJMP2(far, HALT, halve),
    LABEL(halve),
    INSERT_TMACRO(sign_change, CS(left), CS(lo), CS(mid)),
    TCMOV(hi, left, mid, hi),
    TCMOV(lo, left, lo, mid),
    INSERT_TMACRO(midpoint, CS(mid)),
    INSERT_TMACRO(cubic_at, CS(fm), CS(mid)),
    TABS(fm, fm),
    TGE(far, fm, 0.1),
    IADD(i, i, 1),
    ILT(more, i, 256),
    ILAND(far, far, more),
    IPJMP2(far, HALT, halve)
KERNEL_END(TVAR(mid), IVAR(i));
// clang-format on
} // namespace loom_amd::programs
