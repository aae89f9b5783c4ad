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
#include <thread>
#include <vector>

namespace hip = loom::vm::hip;
using loom_amd::programs::collatz;
using loom_amd::programs::factorial;
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

static std::uint64_t
factorial_closed(std::uint64_t a)
{
    std::uint64_t r = 1;
    for(std::uint64_t k = 2; k <= a; ++k) r *= k;
    return r;
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
    // stream bookkeeping: every row crosses PCIe once in and once out, so
    // the bytes moved are proportional to the threads exchanged
    {
        std::vector<std::uint64_t> x(5000);
        for(std::size_t i = 0; i < x.size(); ++i) x[i] = 1 + (i * 40503u) % 70000;
        hip::params p;
        p.live_threshold = 32;
        hip::hybrid_stream<collatz<std::uint64_t>> s(512, p);
        auto out = drive_stream(s, x);
        EXPECT(out.size() == x.size(), "stream size %zu", out.size());
        const loom_stream_counters c = s.counters();
        EXPECT(c.rows_in == x.size() && c.rows_out == x.size(), "rows %llu in, %llu out",
               (unsigned long long)c.rows_in, (unsigned long long)c.rows_out);
        // 8 B of input per row in; 4 B of output + 8 B of sequence number out
        EXPECT(c.bytes_h2d == 8 * x.size(), "bytes in %llu", (unsigned long long)c.bytes_h2d);
        EXPECT(c.bytes_d2h == 12 * x.size(), "bytes out %llu", (unsigned long long)c.bytes_d2h);
        EXPECT(c.exchanges > 1, "exchanges %llu", (unsigned long long)c.exchanges);
        std::vector<std::uint32_t> want(x.size());
        for(std::size_t i = 0; i < x.size(); ++i) want[i] = collatz_steps(x[i]);
        std::ranges::sort(out);
        std::ranges::sort(want);
        EXPECT(out == want, "stream (live threshold 32) outputs differ");
    }
    // several host threads, each with its own execution context, running
    // the same kernel at once: every result equals the closed form
    {
        std::vector<std::thread> ts;
        std::vector<int>         bad(4, 0);
        for(int t = 0; t < 4; ++t)
            ts.emplace_back(
                [t, &bad]
                {
                    loom_context* ctx = nullptr;
                    if(loom_context_create(&ctx) != 0)
                    {
                        bad[t] = -1;
                        return;
                    }
                    const int k = hip::kernel_index<collatz<std::uint64_t>>();
                    std::vector<std::uint64_t> x(20000 + 1000 * t);
                    for(std::size_t i = 0; i < x.size(); ++i) x[i] = 1 + (i * 7u + t) % 500000;
                    std::vector<std::uint32_t> y(x.size());
                    const void*                in[1]  = {x.data()};
                    void*                      out[1] = {y.data()};
                    loom_params                p{};
                    p.live_threshold = 48;
                    p.refill         = 1;
                    p.reconv         = 2;
                    for(int rep = 0; rep < 5; ++rep)
                    {
                        if(loom_ctx_run_host(ctx, k, LOOM_MODE_HYBRID, in, out, x.size(), &p,
                                             nullptr) != 0)
                            bad[t] = -2;
                        for(std::size_t i = 0; i < x.size(); ++i)
                            if(y[i] != collatz_steps(x[i])) ++bad[t];
                    }
                    loom_context_destroy(ctx);
                });
        for(auto& th: ts) th.join();
        for(int t = 0; t < 4; ++t) EXPECT(bad[t] == 0, "context thread %d: %d", t, bad[t]);
    }
    // the Shop: 4 worker threads x 3 kernels, each worker a device-resident
    // stream per kernel, the main thread exchanging (reference
    // src/tests/vm_mp.cpp:90-150); outputs compared as sorted multisets
    {
        std::vector<std::uint64_t> in_fac(3000), in_fib(5000), in_cz(7000);
        for(std::size_t i = 0; i < in_fac.size(); ++i) in_fac[i] = i % 21;
        for(std::size_t i = 0; i < in_fib.size(); ++i) in_fib[i] = (i * 31) % 91;
        for(std::size_t i = 0; i < in_cz.size(); ++i) in_cz[i] = 1 + (i * 2654435761u) % 300000;
        std::vector<std::uint64_t> out_fac, out_fib;
        std::vector<std::uint32_t> out_cz;
        hip::params p;
        p.live_threshold = 32;
        hip::shop<factorial<std::uint64_t>, fib<std::uint64_t>, collatz<std::uint64_t>> shop(
            4, 256, p);
        auto oc = std::make_tuple(std::back_inserter(out_fac), std::back_inserter(out_fib),
                                  std::back_inserter(out_cz));
        auto ic = shop.bind_inputs(in_fac, in_fib, in_cz);
        {
            auto pool = shop.get_thread_pool();
            while(!hip::is_complete(ic)) shop.exchange(ic, oc);
            shop.set();
        }
        shop.finalize(oc);
        std::vector<std::uint64_t> w_fac, w_fib;
        std::vector<std::uint32_t> w_cz;
        for(auto a: in_fac) w_fac.push_back(factorial_closed(a));
        for(auto n: in_fib) w_fib.push_back(fib_closed(n));
        for(auto x: in_cz) w_cz.push_back(collatz_steps(x));
        for(auto* v: {&out_fac, &out_fib, &w_fac, &w_fib}) std::ranges::sort(*v);
        std::ranges::sort(out_cz);
        std::ranges::sort(w_cz);
        EXPECT(out_fac == w_fac, "shop factorial outputs differ (%zu)", out_fac.size());
        EXPECT(out_fib == w_fib, "shop fib outputs differ (%zu)", out_fib.size());
        EXPECT(out_cz == w_cz, "shop collatz outputs differ (%zu)", out_cz.size());
    }
    std::printf("facade_test: %s (%d failures)\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
