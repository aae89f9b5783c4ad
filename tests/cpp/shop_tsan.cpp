// The Shop's coordination (include/loom_vm/vm/hip.hpp basic_shop and
// basic_context, the reference's mp::Shop / mp::Context protocol,
// src/vm/mp.hpp:61-385) over a CPU stand-in stream, built with
// ThreadSanitizer by tests/test_shop_tsan.py: worker threads step and sort
// while the main thread exchanges, as on the GPU, and every output must come
// back exactly once.  The stand-in stream serialises its calls with a lock,
// as the device-resident stream of the C ABI does.
#include <vm/hip.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

namespace hip = loom::vm::hip;

// "kernels": a thread of input x needs 1 + x % 5 steps and outputs f(x)
struct Square
{
    static std::uint64_t f(std::uint64_t x) { return x * x; }
};
struct Triple
{
    static std::uint64_t f(std::uint64_t x) { return 3 * x + 1; }
};

template <typename K>
class cpu_stream
{
    struct Slot
    {
        std::uint64_t x = 0;
        int           left = 0;  // steps to HALT
        int           state = 0; // 0 empty, 1 live, 2 halted
    };
    mutable std::mutex m_mu;
    std::vector<Slot>  m_slots;

  public:
    using program = K;

    explicit cpu_stream(std::size_t n, int = 0) : m_slots(n) {}

    template <typename R>
    auto
    set_inputs(R&& input)
    {
        std::lock_guard<std::mutex> lk(m_mu);
        auto it = std::ranges::begin(input);
        for(auto& s: m_slots)
        {
            s = Slot{};
            if(it != std::ranges::end(input)) s = Slot{*it, 1 + static_cast<int>(*it % 5), 1}, ++it;
        }
        return it;
    }

    void
    operator()()
    {
        std::lock_guard<std::mutex> lk(m_mu);
        for(auto& s: m_slots)
            if(s.state == 1 && --s.left == 0) s.state = 2;
    }

    bool
    sort()
    {
        std::lock_guard<std::mutex> lk(m_mu);
        return std::ranges::any_of(m_slots, [](const Slot& s) { return s.state == 2; });
    }

    template <typename R>
    std::vector<std::uint64_t>
    exchange(R&& input)
    {
        std::lock_guard<std::mutex> lk(m_mu);
        std::vector<std::uint64_t> out;
        auto                       it = std::ranges::begin(input);
        for(auto& s: m_slots)
            if(s.state == 2 && it != std::ranges::end(input))
            {
                out.push_back(K::f(s.x));
                s = Slot{*it, 1 + static_cast<int>(*it % 5), 1};
                ++it;
            }
        return out;
    }

    bool
    done() const
    {
        std::lock_guard<std::mutex> lk(m_mu);
        return std::ranges::none_of(m_slots, [](const Slot& s) { return s.state == 1; });
    }

    std::vector<std::uint64_t>
    get_output_zipped()
    {
        std::lock_guard<std::mutex> lk(m_mu);
        std::vector<std::uint64_t> out;
        for(auto& s: m_slots)
            if(s.state == 2)
            {
                out.push_back(K::f(s.x));
                s.state = 0;
            }
        return out;
    }
};

int
main()
{
    int failures = 0;
    for(std::size_t workers: {1u, 3u, 8u})
    {
        std::vector<std::uint64_t> a(20000), b(13000);
        for(std::size_t i = 0; i < a.size(); ++i) a[i] = i;
        for(std::size_t i = 0; i < b.size(); ++i) b[i] = 7 * i + 3;
        std::vector<std::uint64_t> oa, ob;
        hip::basic_shop<cpu_stream, Square, Triple> shop(workers, 64, 0);
        auto oc = std::make_tuple(std::back_inserter(oa), std::back_inserter(ob));
        auto ic = shop.bind_inputs(a, b);
        {
            auto pool = shop.get_thread_pool();
            while(!hip::is_complete(ic)) shop.exchange(ic, oc);
            shop.set();
        }
        shop.finalize(oc);
        std::vector<std::uint64_t> wa, wb;
        for(auto x: a) wa.push_back(Square::f(x));
        for(auto x: b) wb.push_back(Triple::f(x));
        for(auto* v: {&oa, &ob, &wa, &wb}) std::ranges::sort(*v);
        if(oa != wa || ob != wb)
        {
            std::fprintf(stderr, "FAIL: %zu workers: %zu/%zu and %zu/%zu outputs\n", workers,
                         oa.size(), wa.size(), ob.size(), wb.size());
            ++failures;
        }
    }
    std::printf("shop_tsan: %s\n", failures ? "FAIL" : "ok");
    return failures ? 1 : 0;
}
