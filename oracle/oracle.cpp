// -*- C++ -*-
/// @file oracle.cpp
/// @brief ORACLE (test infrastructure only): C entry points of the scalar
/// restatement, for the parity tests and the CPU baseline timing.  Never
/// linked into or called by the product library.
#include "bse.hpp"
#include "smallpt9_direct.hpp"

#ifdef LOOM_ORACLE_REF
// the reference's own kernels (oracle/_ref/liboracle_ref.so, built by build_ref.py)
#include "ref_kernels.hpp"
#define LOOM_ORACLE_CATALOG LOOM_REF_CATALOG
#else
#include "../loom_amd/kernels/catalog.hpp"
#define LOOM_ORACLE_CATALOG LOOM_AMD_CATALOG
#endif

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

using namespace loom_oracle;

namespace
{
struct Entry
{
    const char* name;
    Program (*build)();
};

#define LOOM_ORACLE_ENTRY(NAME, ...) Entry{NAME, &build_program<__VA_ARGS__>},
const Entry g_entries[] = {LOOM_ORACLE_CATALOG(LOOM_ORACLE_ENTRY)};
#undef LOOM_ORACLE_ENTRY

constexpr int g_count = sizeof(g_entries) / sizeof(g_entries[0]);

std::vector<Program>&
programs()
{
    static std::vector<Program> p = []
    {
        std::vector<Program> v;
        for(const auto& e: g_entries) v.push_back(e.build());
        return v;
    }();
    return p;
}

void
run_range(const Program& p, const void* const* in, void* const* out, std::uint64_t begin,
          std::uint64_t end, std::atomic<std::uint64_t>& executed, std::atomic<std::uint32_t>& err)
{
    std::vector<std::uint64_t> R(p.reg_name.size());
    std::uint64_t              ex = 0;
    std::uint32_t              e  = 0;
    for(std::uint64_t t = begin; t < end; ++t)
    {
        std::fill(R.begin(), R.end(), 0);
        for(std::size_t k = 0; k < p.in_reg.size(); ++k)
        {
            const int b = ty_bytes(p.in_ty[k]);
            std::uint64_t v = 0;
            std::memcpy(&v, static_cast<const char*>(in[k]) + t * b, b);
            R[p.in_reg[k]] = v;
        }
        ex += run_thread(p, R.data(), e);
        for(std::size_t k = 0; k < p.out_reg.size(); ++k)
        {
            const int b = ty_bytes(p.out_ty[k]);
            std::memcpy(static_cast<char*>(out[k]) + t * b, &R[p.out_reg[k]], b);
        }
    }
    executed += ex;
    err |= e;
}
} // namespace

extern "C" {

int
oracle_kernel_count(void)
{
    return g_count;
}

const char*
oracle_kernel_name(int k)
{
    return k >= 0 && k < g_count ? g_entries[k].name : nullptr;
}

int
oracle_kernel_find(const char* name)
{
    for(int k = 0; k < g_count; ++k)
        if(std::strcmp(g_entries[k].name, name) == 0) return k;
    return -1;
}

/// Types of inputs/outputs (codes as loom_type) and register count.
int
oracle_kernel_io(int k, int* nin, int* nout, int* in_types, int* out_types, int* nregs)
{
    if(k < 0 || k >= g_count) return 1;
    const Program& p = programs()[k];
    *nin             = static_cast<int>(p.in_ty.size());
    *nout            = static_cast<int>(p.out_ty.size());
    for(std::size_t i = 0; i < p.in_ty.size(); ++i) in_types[i] = static_cast<int>(p.in_ty[i]);
    for(std::size_t i = 0; i < p.out_ty.size(); ++i) out_types[i] = static_cast<int>(p.out_ty[i]);
    if(nregs) *nregs = static_cast<int>(p.reg_name.size());
    return 0;
}

/// Run threads [begin, begin+count) of host columns; `threads` CPU workers.
int
oracle_run(int k, const void* const* in, void* const* out, uint64_t begin, uint64_t count,
           int threads, uint64_t* executed, uint32_t* errors)
{
    if(k < 0 || k >= g_count) return 1;
    const Program&             p = programs()[k];
    std::atomic<std::uint64_t> ex{0};
    std::atomic<std::uint32_t> er{0};
    threads = std::max(1, threads);
    if(threads == 1 || count < 1024)
        run_range(p, in, out, begin, begin + count, ex, er);
    else
    {
        std::vector<std::thread> pool;
        const std::uint64_t      chunk = (count + threads - 1) / threads;
        for(int t = 0; t < threads; ++t)
        {
            const std::uint64_t b = begin + chunk * t;
            const std::uint64_t e = std::min(begin + count, b + chunk);
            if(b >= e) break;
            pool.emplace_back([&, b, e] { run_range(p, in, out, b, e, ex, er); });
        }
        for(auto& th: pool) th.join();
    }
    if(executed) *executed = ex;
    if(errors) *errors = er;
    return 0;
}

/// Work of threads [begin, begin+count), counted the way the reference
/// prices a block (simd.hpp:2341-2665 use |B| = #Compute + 1): per executed
/// block one thread step, #Compute + 1 ops, and one flop per floating-point
/// Compute other than moves, selects, table reads and checks (two for
/// fma/fms).  `threads` CPU workers.
int
oracle_work(int k, const void* const* in, uint64_t begin, uint64_t count, int threads,
            uint64_t* steps, uint64_t* ops, uint64_t* flops)
{
    if(k < 0 || k >= g_count || !in || !steps || !ops || !flops) return 1;
    const Program& p = programs()[k];
    // cost of the block starting at each statement: computes up to the next label
    std::vector<std::uint32_t> bops(p.code.size(), 0), bflops(p.code.size(), 0);
    for(std::size_t i = 0; i < p.code.size(); ++i)
    {
        if(p.code[i].kind == Stmt::label) continue;
        std::uint32_t nc = 0, nf = 0;
        for(std::size_t j = i; j < p.code.size() && p.code[j].kind != Stmt::label; ++j)
        {
            const Stmt& s = p.code[j];
            if(s.kind != Stmt::compute) continue;
            ++nc;
            const bool fp = s.ety == Ty::f32 || s.ety == Ty::f64;
            if(fp && s.code != Code::mov && s.code != Code::cmov && s.code != Code::sct &&
               s.code != Code::slt && s.code != Code::assert_fp)
                nf += (s.code == Code::fma || s.code == Code::fms) ? 2 : 1;
        }
        bops[i]   = nc + 1;
        bflops[i] = nf;
    }
    std::atomic<std::uint64_t> st{0}, op{0}, fl{0};
    auto range = [&](std::uint64_t b, std::uint64_t e)
    {
        std::vector<std::uint64_t> R(p.reg_name.size());
        std::uint64_t              ls = 0, lo = 0, lf = 0;
        std::uint32_t              err = 0;
        for(std::uint64_t t = b; t < e; ++t)
        {
            std::fill(R.begin(), R.end(), 0);
            for(std::size_t j = 0; j < p.in_reg.size(); ++j)
            {
                const int     w = ty_bytes(p.in_ty[j]);
                std::uint64_t v = 0;
                std::memcpy(&v, static_cast<const char*>(in[j]) + t * w, w);
                R[p.in_reg[j]] = v;
            }
            run_blocks(p, R.data(), err, ~std::uint64_t(0),
                       [&](std::size_t f)
                       {
                           ++ls;
                           lo += bops[f];
                           lf += bflops[f];
                           return true;
                       });
        }
        st += ls;
        op += lo;
        fl += lf;
    };
    threads = std::max(1, threads);
    if(threads == 1 || count < 1024)
        range(begin, begin + count);
    else
    {
        std::vector<std::thread> pool;
        const std::uint64_t      chunk = (count + threads - 1) / threads;
        for(int w = 0; w < threads; ++w)
        {
            const std::uint64_t b = begin + chunk * w;
            const std::uint64_t e = std::min(begin + count, b + chunk);
            if(b >= e) break;
            pool.emplace_back(range, b, e);
        }
        for(auto& th: pool) th.join();
    }
    *steps = st;
    *ops   = op;
    *flops = fl;
    return 0;
}

/// Step trace of thread t of host columns: after every block, first[s] = the
/// index of the block's first statement, errs[s] = error bits so far and
/// regs[s * nregs + r] = bits of register r (oracle register order).
int
oracle_trace(int k, const void* const* in, uint64_t t, uint32_t max_steps, int32_t* first,
             uint64_t* regs, uint32_t* errs, uint32_t* steps)
{
    if(k < 0 || k >= g_count || !in || !first || !regs || !errs || !steps) return 1;
    const Program&             p = programs()[k];
    const std::size_t          nr = p.reg_name.size();
    std::vector<std::uint64_t> R(nr, 0);
    for(std::size_t j = 0; j < p.in_reg.size(); ++j)
    {
        const int     b = ty_bytes(p.in_ty[j]);
        std::uint64_t v = 0;
        std::memcpy(&v, static_cast<const char*>(in[j]) + t * b, b);
        R[p.in_reg[j]] = v;
    }
    std::uint32_t e = 0, s = 0;
    run_blocks(p, R.data(), e, ~std::uint64_t(0),
               [&](std::size_t f)
               {
                   first[s] = static_cast<int32_t>(f);
                   errs[s]  = e;
                   std::copy(R.begin(), R.end(), regs + std::size_t(s) * nr);
                   return ++s < max_steps;
               });
    *steps = s;
    return 0;
}

/// Space-separated names of the labels that start at statement `stmt`
/// ("START" first for the first statement).
int
oracle_block_labels(int k, int32_t stmt, char* buf, size_t len)
{
    if(k < 0 || k >= g_count || !buf || !len) return 1;
    const Program& p = programs()[k];
    if(stmt < 0 || static_cast<std::size_t>(stmt) >= p.code.size()) return 1;
    const std::string l = labels_before(p, static_cast<std::size_t>(stmt));
    const std::size_t n = std::min(len - 1, l.size());
    std::memcpy(buf, l.data(), n);
    buf[n] = 0;
    return 0;
}

/// Name of register r (oracle register order).
int
oracle_register_name(int k, int r, char* buf, size_t len)
{
    if(k < 0 || k >= g_count || !buf || !len) return 1;
    const Program& p = programs()[k];
    if(r < 0 || static_cast<std::size_t>(r) >= p.reg_name.size()) return 1;
    const std::string& l = p.reg_name[r];
    const std::size_t  n = std::min(len - 1, l.size());
    std::memcpy(buf, l.data(), n);
    buf[n] = 0;
    return 0;
}

/// Literal operand i of kernel k, in statement order: its text (NUL-terminated,
/// at most len bytes), type code (loom_type) and the bit pattern the kernel
/// computes with.  Returns 1 past the last literal.
int
oracle_kernel_literal(int k, int i, char* text, size_t len, int* type, uint64_t* bits)
{
    if(k < 0 || k >= g_count || i < 0 || !text || !len) return 1;
    const Program& p = programs()[k];
    int            n = 0;
    for(const auto& s: p.code)
        for(const auto& a: s.args)
        {
            if(!a.literal || n++ != i) continue;
            const std::size_t m = std::min(len - 1, a.name.size());
            std::memcpy(text, a.name.data(), m);
            text[m] = 0;
            *type   = static_cast<int>(a.ty);
            *bits   = a.bits;
            return 0;
        }
    return 1;
}

int
oracle_parse_f32(const char* s, float* v)
{
    auto r = lit::parse<float>(s);
    if(!r) return 1;
    *v = *r;
    return 0;
}

int
oracle_parse_f64(const char* s, double* v)
{
    auto r = lit::parse<double>(s);
    if(!r) return 1;
    *v = *r;
    return 0;
}

int
oracle_parse_u64(const char* s, uint64_t* v)
{
    auto r = lit::parse<std::uint64_t>(s);
    if(!r) return 1;
    *v = *r;
    return 0;
}

int
oracle_parse_i32(const char* s, int32_t* v)
{
    auto r = lit::parse<std::int32_t>(s);
    if(!r) return 1;
    *v = *r;
    return 0;
}

/// The direct C++ path tracer on samples [begin, begin+count) of host columns.
int
oracle_smallpt9_direct(const uint32_t* sid, const uint32_t* job, float* lx, float* ly, float* lz,
                       uint64_t begin, uint64_t count, int threads)
{
    static const direct::Scene scene;
    auto work = [&](std::uint64_t b, std::uint64_t e)
    {
        for(std::uint64_t i = b; i < e; ++i) direct::sample(scene, sid[i], job[i], lx[i], ly[i], lz[i]);
    };
    threads = std::max(1, threads);
    if(threads == 1 || count < 1024)
        work(begin, begin + count);
    else
    {
        std::vector<std::thread> pool;
        const std::uint64_t      chunk = (count + threads - 1) / threads;
        for(int t = 0; t < threads; ++t)
        {
            const std::uint64_t b = begin + chunk * t;
            const std::uint64_t e = std::min(begin + count, b + chunk);
            if(b >= e) break;
            pool.emplace_back(work, b, e);
        }
        for(auto& th: pool) th.join();
    }
    return 0;
}

/// Collatz total stopping time by its definition (x -> x/2 if even, 3x+1
/// if odd, counted until x == 1; 0 for x <= 1), written directly in C++ and
/// independent of the IR: the brute force the Collatz kernel is checked
/// against.  Returns 1 if a trajectory would leave 64 bits.
int
oracle_collatz_direct(const uint64_t* x, uint32_t* steps, uint64_t count, int threads)
{
    std::atomic<int> overflow{0};
    auto work = [&](std::uint64_t b, std::uint64_t e)
    {
        for(std::uint64_t i = b; i < e; ++i)
        {
            std::uint64_t v = x[i];
            std::uint32_t n = 0;
            while(v > 1)
            {
                if(v & 1)
                {
                    if(v > (~std::uint64_t(0) - 1) / 3) overflow = 1;
                    v = 3 * v + 1;
                }
                else
                    v >>= 1;
                ++n;
            }
            steps[i] = n;
        }
    };
    threads = std::max(1, threads);
    std::vector<std::thread> pool;
    const std::uint64_t      chunk = (count + threads - 1) / threads;
    for(int t = 0; t < threads; ++t)
    {
        const std::uint64_t b = chunk * t;
        const std::uint64_t e = std::min(count, b + chunk);
        if(b >= e) break;
        pool.emplace_back(work, b, e);
    }
    for(auto& th: pool) th.join();
    return overflow.load();
}

float
oracle_sin_f32(float x)
{
    return sin_poly(x);
}

float
oracle_cos_f32(float x)
{
    return cos_poly(x);
}

double
oracle_sin_f64(double x)
{
    return sin_poly(x);
}

double
oracle_cos_f64(double x)
{
    return cos_poly(x);
}
}
