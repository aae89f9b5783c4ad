// -*- C++ -*-
/// @file dump.hpp
/// @brief A text listing of a lowered kernel: blocks with their cost, live-in
/// registers and successors, and every instruction with its operands — the
/// block/instruction/register view of the reference's IR inspector
/// (dbg.hpp:1183-1202), produced from the lowered table (host code).
#pragma once

#include <vm/lower.hpp>

#include <cstdio>
#include <string>

namespace loom::vm::ir
{
inline const char*
ty_name(Ty t)
{
    static constexpr const char* n[] = {"u8", "u16", "u32", "u64", "i8",
                                        "i16", "i32", "i64", "f32", "f64", "-"};
    return n[static_cast<int>(t) <= static_cast<int>(Ty::none) ? static_cast<int>(t) : 10];
}

inline const char*
op_name(Op o)
{
    static constexpr const char* n[] = {
        "mov",  "add",   "sub",   "mul",   "div",       "mod",  "neg",   "eq",    "ne",
        "lt",   "gt",    "le",    "ge",    "land",      "lor",  "lnot",  "band",  "bor",
        "bxor", "bnot",  "shl",   "shr",   "cmov",      "abs",  "fma",   "fms",   "rsqrt",
        "sct",  "slt",   "pacc",  "assert_fp", "min",   "max",  "round", "ceil",  "floor",
        "trunc", "sin",  "cos",   "sqrt",  "copysign",  "-"};
    const int i = static_cast<int>(o);
    return i < static_cast<int>(sizeof(n) / sizeof(n[0])) ? n[i] : "?";
}

template <typename Tbl>
std::string
dump_table(const Tbl& t)
{
    std::string s;
    char        line[256];
    auto        put = [&](const char* fmt, auto... a)
    {
        std::snprintf(line, sizeof(line), fmt, a...);
        s += line;
    };
    auto label = [&](std::uint16_t b) -> std::string
    {
        if(b == halt16) return "HALT";
        const std::string_view l = t.block_label[b];
        return l.empty() ? (b == 0 ? std::string("START") : "b" + std::to_string(b)) : std::string(l);
    };
    auto operand = [&](std::uint16_t o) -> std::string
    {
        const Operand& x = t.opnd[o];
        if(!x.lit) return std::string(x.name);
        return std::string(x.name) + "'";
    };

    put("kernel %.*s  hash 0x%016llx  index %s\n", static_cast<int>(t.kname.size()), t.kname.data(),
        static_cast<unsigned long long>(t.hash), t.index_bytes == 8 ? "u64" : (t.index_bytes == 2 ? "u16" : "u32"));
    put("%zu blocks, %zu instructions, %zu registers, state %zu B\n", t.nb, t.ni, t.nr, t.state_size);
    s += "inputs:";
    for(std::size_t i = 0; i < t.nin; ++i)
        put(" %.*s:%s", static_cast<int>(t.reg[t.in_slot[i]].name.size()),
            t.reg[t.in_slot[i]].name.data(), ty_name(t.in_ty[i]));
    s += "\noutputs:";
    for(std::size_t i = 0; i < t.nout; ++i)
        put(" %.*s:%s", static_cast<int>(t.reg[t.out_slot[i]].name.size()),
            t.reg[t.out_slot[i]].name.data(), ty_name(t.out_ty[i]));
    s += "\n";
    for(std::size_t b = 0; b < t.nb; ++b)
    {
        const Block& k = t.block[b];
        put("\nblock %zu %s  ops %u flops %u  rank %u\n", b, label(static_cast<std::uint16_t>(b)).c_str(),
            k.ops, k.flops, t.rank[b]);
        s += "  live-in:";
        for(std::size_t r = 0; r < t.nr; ++r)
            if(k.live_in.test(r)) put(" %.*s", static_cast<int>(t.reg[r].name.size()), t.reg[r].name.data());
        s += "\n  successors:";
        for(std::size_t j = 0; j < k.ns; ++j) put(" %s", label(t.succ[k.s0 + j]).c_str());
        s += "\n";
        for(std::size_t i = k.i0; i < k.i0 + k.ni; ++i)
        {
            const Instr& in = t.instr[i];
            if(in.kind == SK::jump)
                put("    jump %s\n", label(t.target[in.t0]).c_str());
            else if(in.kind == SK::pjump)
            {
                put("    pjump %s ->", operand(in.a0).c_str());
                for(std::size_t j = 0; j < in.nt; ++j) put(" %s", label(t.target[in.t0 + j]).c_str());
                s += "\n";
            }
            else
            {
                put("    %s = %s.%s", operand(in.dst).c_str(), op_name(in.op), ty_name(in.ety));
                for(std::size_t j = 0; j < in.na; ++j) put(" %s", operand(static_cast<std::uint16_t>(in.a0 + j)).c_str());
                if(in.nc) put(" [%u constants]", in.nc);
                s += "\n";
            }
        }
    }
    return s;
}
} // namespace loom::vm::ir
