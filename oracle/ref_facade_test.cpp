// -*- C++ -*-
/// @file ref_facade_test.cpp
/// @brief TEST INFRASTRUCTURE: the reference's own test programs
/// (src/tests/programs.hpp) driven through this repository's C++ facade
/// (include/loom_vm/vm/hip.hpp) the way the reference's CUDA backend test
/// drives them (src/tests/vm.cu:58-86: a function per program and a stream
/// for factorial), checked against the reference's own validators
/// (programs.hpp:288-391).  The backend type is the only change from the
/// reference's call shapes.  Built by oracle/build_ref.py into oracle/_ref/
/// (it links libloom_refkernels.so, which registers the programs' gfx950
/// code); exit status 0 = pass.
#include <vm/hip.hpp>

#include <common/math.hpp> // loom::fac, for val_fac
#include <tests/programs.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <ranges>
#include <vector>

namespace hip = loom::vm::hip;

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

/// Inputs drawn from [lo, hi] by a fixed multiplicative sequence.
template <typename T>
static std::vector<T>
spread(std::size_t n, T lo, T hi)
{
    std::vector<T> v(n);
    const std::uint64_t span = static_cast<std::uint64_t>(hi - lo) + 1;
    for(std::size_t i = 0; i < n; ++i)
        v[i] = static_cast<T>(lo + static_cast<T>((i * 2654435761ull + 12345) % span));
    return v;
}

/// Each form of a one-input, one-output program against validator @p val.
template <typename K, typename In, typename Val>
static void
check_function(const char* name, const std::vector<In>& in, Val val)
{
    {
        hip::compact_function<K> f;
        auto out = f.template operator()<true>(in);
        EXPECT(out.size() == in.size(), "%s compact: %zu outputs", name, out.size());
        for(std::size_t i = 0; i < in.size() && i < out.size(); ++i)
            EXPECT(out[i] == val(in[i]), "%s compact, input %zu", name, i);
    }
    {
        hip::hybrid_function<K> f;
        auto out = f.template operator()<true>(in);
        EXPECT(out.size() == in.size(), "%s hybrid: %zu outputs", name, out.size());
        for(std::size_t i = 0; i < in.size() && i < out.size(); ++i)
            EXPECT(out[i] == val(in[i]), "%s hybrid, input %zu", name, i);
    }
}

/// The reference's stream protocol (src/tests/stream_wrapper.hpp:33-80):
/// set_inputs, step, exchange while input remains, step until done,
/// collect.
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
    // the programs and generator ranges of the reference's CUDA test
    // (src/tests/vm.cu:31-41, 52-58), at a ragged size
    const std::size_t n = 4099;
    check_function<increment<int>>("increment<int>", spread<int>(n, 2, 10), val_inc);
    check_function<fac_test<unsigned>>("fac_test<unsigned>", spread<unsigned>(n, 0, 10), val_fac);
    check_function<fac_test<std::uint64_t>>("fac_test<u64>", spread<std::uint64_t>(n, 0, 20),
                                            val_fac);
    check_function<fib_test<std::uint64_t>>("fib_test<u64>", spread<std::uint64_t>(n, 0, 80),
                                            val_fib);
    check_function<jmp4_test<float>>("jmp4_test<float>", spread<unsigned>(n, 0, 3), val_jmp4);
    {
        std::vector<float> x(n);
        for(std::size_t i = 0; i < n; ++i) x[i] = static_cast<float>(i) / static_cast<float>(n);
        check_function<union_test<float>>("union_test<float>", x, val_union1);
        check_function<union_cmov_test<float>>("union_cmov_test<float>", x, val_union2);
    }
    // the stream test of the reference's CUDA backend: factorial through a
    // stream of half the input's size (src/tests/vm.cu:48-50, 79-82),
    // outputs compared as multisets (the stream returns them unordered)
    {
        const auto                x = spread<unsigned>(n, 0, 10);
        hip::hybrid_stream<fac_test<unsigned>> s(n / 2);
        auto                      out = drive_stream(s, x);
        std::vector<unsigned>     want(x.size());
        for(std::size_t i = 0; i < x.size(); ++i) want[i] = val_fac(x[i]);
        EXPECT(out.size() == want.size(), "stream fac_test: %zu outputs", out.size());
        std::ranges::sort(out);
        std::ranges::sort(want);
        EXPECT(out == want, "stream fac_test outputs differ");
    }
    std::printf("ref_facade_test: %s (%d failures)\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
