// Host program driving the gfx950 backend through the C++ facade
// (include/loom_vm/vm/hip.hpp) the way the reference's VM tests drive its
// SIMD backends (reference src/tests/vm.cpp:40-260, stream protocol of
// src/tests/stream_wrapper.hpp:33-80).  Expected values come from closed
// forms and brute force computed here on the host.  Exit status 0 = pass.
#include <vm/hip.hpp>

#include "../../loom_amd/kernels/programs.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <vector>

namespace hip = loom::vm::hip;
using loom_amd::programs::collatz;
using loom_amd::programs::fib;

static int failures = 0;

#define EXPECT(cond, ...)                                                                          \
    do                                                                                             \
    {                                                                                              \
        if(!(cond))                                                                                \
        {                                                                                          \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                              \
            std::fprintf(stderr, __VA_ARGS__);                                                     \
            std::fprintf(stderr, "\n");                                                            \
            ++failures;                                                                            \
        }                                                                                          \
    } while(0)

static std::uint64_t
fib_closed(std::uint64_t n)
{
    // iterative closed recurrence modulo 2^64
    std::uint64_t a = 0, b = 1;
    for(std::uint64_t i = 0; i < n; ++i)
    {
        const std::uint64_t c = a + b;
        a                     = b;
        b                     = c;
    }
    return a;
}

static std::uint32_t
collatz_steps(std::uint64_t x)
{
    std::uint32_t s = 0;
    while(x > 1)
    {
        x = (x & 1) ? 3 * x + 1 : x / 2;
        ++s;
    }
    return s;
}

// stream driver with the reference protocol: set_inputs, step, exchange
// until the input is consumed, then step until done, then collect
template <typename S, typename In>
static auto
drive_stream(S& s, const In& input)
{
    std::vector<hip::output_type<typename S::program>> out;
    auto step = [&]
    {
        s();
        s.sort();
    };
    auto i = s.set_inputs(input);
    step();
    while(i != input.end())
    {
        auto ex = s.exchange(std::ranges::subrange(i, input.end()));
        i += static_cast<std::ptrdiff_t>(ex.size());
        out.insert(out.end(), ex.begin(), ex.end());
        step();
    }
    while(!s.done()) step();
    auto rest = s.get_output_zipped();
    out.insert(out.end(), rest.begin(), rest.end());
    return out;
}

int
main()
{
    // fib, compact and hybrid, n = 0..92 repeated (ragged: 4099 threads)
    std::vector<std::uint64_t> n(4099);
    for(std::size_t i = 0; i < n.size(); ++i) n[i] = (i * 7919) % 93;
    {
        hip::compact_function<fib<std::uint64_t>> f;
        auto out = f.template operator()<true>(n);
        EXPECT(out.size() == n.size(), "compact fib size %zu", out.size());
        for(std::size_t i = 0; i < n.size(); ++i)
            EXPECT(out[i] == fib_closed(n[i]), "compact fib(%llu)", (unsigned long long)n[i]);
    }
    {
        hip::hybrid_function<fib<std::uint64_t>> f;
        auto out = f.template operator()<true>(n);
        for(std::size_t i = 0; i < n.size(); ++i)
            EXPECT(out[i] == fib_closed(n[i]), "hybrid fib(%llu)", (unsigned long long)n[i]);
        EXPECT(f.stats().counters[LOOM_C_SLOT_OPS] >= f.stats().counters[LOOM_C_ACTIVE_OPS],
               "counters");
    }
    // collatz (u64 state, u32 step count), hybrid with queue regrouping
    {
        std::vector<std::uint64_t> x(10007);
        std::iota(x.begin(), x.end(), 1);
        hip::params p;
        p.regroup        = true;
        p.live_threshold = 32;
        hip::hybrid_function<collatz<std::uint64_t>> f(0, p);
        auto out = f.template operator()<true>(x);
        for(std::size_t i = 0; i < x.size(); ++i)
            EXPECT(out[i] == collatz_steps(x[i]), "collatz(%llu) = %u", (unsigned long long)x[i],
                   (unsigned)out[i]);
    }
    // stream: 1000 slots over 25000 inputs, outputs compared as multisets
    {
        std::vector<std::uint64_t> x(25000);
        for(std::size_t i = 0; i < x.size(); ++i) x[i] = 1 + (i * 2654435761u) % 100000;
        hip::hybrid_stream<collatz<std::uint64_t>> s(1000);
        auto out = drive_stream(s, x);
        std::vector<std::uint32_t> want(x.size());
        for(std::size_t i = 0; i < x.size(); ++i) want[i] = collatz_steps(x[i]);
        EXPECT(out.size() == want.size(), "stream size %zu", out.size());
        std::ranges::sort(out);
        std::ranges::sort(want);
        EXPECT(out == want, "stream outputs differ");
    }
    std::printf("facade_test: %s (%d failures)\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
