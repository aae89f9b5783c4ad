// -*- C++ -*-
/// @file hip.hpp
/// @brief The gfx950 backend behind the reference backend interface:
/// `loom::vm::hip::compact_function<K>`, `hybrid_function<K>` and
/// `hybrid_stream<K>`, drop-in for `simd::compact_function<K>` /
/// `simd::hybrid_function<K>` / `simd::hybrid_stream<K>`
/// (reference src/vm/simd.hpp:4310-4335, facade bve.hpp:1028-1270,
/// simd.hpp:2797-2903).
///
/// The device code of K must be compiled into the kernel library (or a
/// module linked with it) with LOOM_HIP_REGISTER; the facade finds it by the
/// structural hash of K's lowered table, so a kernel compiled differently
/// than the caller sees it is rejected.  Every GPU call goes through the C
/// ABI of loom_amd.h.  Like the reference, the facade is noexcept and aborts
/// with a message on a failed call (the reference asserts, cuda.hpp:34-72).
#pragma once

#include <loom_amd.h>

#include <vm/core.hpp>
#include <vm/lower.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <memory>
#include <mutex>
#include <ranges>
#include <thread>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace loom::vm::hip
{
/// @brief Backend parameters (see loom_params).
struct params
{
    std::uint32_t threshold      = 0;
    std::uint32_t live_threshold = 48;
    bool          refill         = true;
    /// hybrid: LOOM_REGROUP_NONE (false), LOOM_REGROUP_QUEUES (true) or
    /// LOOM_REGROUP_WORKGROUP
    std::uint32_t regroup        = LOOM_REGROUP_NONE;
    int           reconv         = -1; ///< -1: per mode default
    std::uint32_t waves_per_cu   = 0;
    std::uint64_t queue_capacity = 0;
    std::uint64_t max_steps      = 0;
    /// compact_function: keep the registers in device memory instead of
    /// VGPRs (LOOM_REGFILE_MEMORY; runs under the sweep policy)
    std::uint32_t register_file  = LOOM_REGFILE_VGPR;
    /// LOOM_REGROUP_WORKGROUP: dispatches between two sorts (0: 64) and
    /// waves per workgroup (16 or 8; 0: 16)
    std::uint32_t regroup_phase  = 0;
    std::uint32_t regroup_waves  = 0;
};

namespace detail
{
[[noreturn]] inline void
die(const char* what, int code)
{
    std::fprintf(stderr, "loom::vm::hip: %s failed (%d): %s\n", what, code, loom_last_error());
    std::abort();
}

inline void
check(int code, const char* what)
{
    if(code != 0) die(what, code);
}

inline void
ensure_init()
{
    static const bool ok = []
    {
        const int r = loom_init(0);
        if(r != 0 && r != LOOM_E_ARG) die("loom_init", r);
        return true;
    }();
    (void)ok;
}

inline loom_params
to_c(const params& p, int mode)
{
    loom_params c{};
    c.threshold      = p.threshold;
    c.live_threshold = mode == LOOM_MODE_HYBRID ? p.live_threshold : 0;
    c.refill         = p.refill;
    c.regroup        = p.regroup;
    c.register_file  = mode == LOOM_MODE_COMPACT ? p.register_file : LOOM_REGFILE_VGPR;
    c.reconv         = p.reconv >= 0 ? static_cast<std::uint32_t>(p.reconv) :
                       (mode == LOOM_MODE_HYBRID || c.register_file == LOOM_REGFILE_MEMORY ? 2u : 0u);
    c.waves_per_cu   = p.waves_per_cu;
    c.queue_capacity = p.queue_capacity;
    c.max_steps      = p.max_steps;
    c.regroup_phase  = p.regroup_phase;
    c.regroup_waves  = p.regroup_waves;
    return c;
}

template <typename L>
struct types_of;
template <typename... Vs>
struct types_of<loom::type_sequence<Vs...>>
{
    using tuple = std::tuple<vartype<Vs>...>;
    static constexpr std::size_t size = sizeof...(Vs);
};

template <typename T>
concept Column = std::ranges::sized_range<T> && !std::is_arithmetic_v<std::remove_cvref_t<T>>;
} // namespace detail

/// @brief Registered kernel index of K (found by structural hash).
template <Kernel K>
int
kernel_index()
{
    static const int idx = []
    {
        detail::ensure_init();
        constexpr std::uint64_t h = ir::table_of<K>.hash;
        for(int i = 0; i < loom_kernel_count(); ++i)
        {
            loom_kernel_info info{};
            detail::check(loom_kernel_info_get(i, &info), "loom_kernel_info_get");
            if(info.hash == h) return i;
        }
        std::fprintf(stderr, "loom::vm::hip: kernel %.*s is not compiled into the library\n",
                     static_cast<int>(K::nameview.size()), K::nameview.data());
        std::abort();
    }();
    return idx;
}

template <Kernel K>
using input_tuple = typename detail::types_of<typename K::input_vars>::tuple;
template <Kernel K>
using output_tuple = typename detail::types_of<typename K::output_vars>::tuple;

namespace detail
{
template <typename Tuple>
struct maybe_single
{
    using type = Tuple;
};
template <typename T>
struct maybe_single<std::tuple<T>>
{
    using type = T;
};
} // namespace detail

/// @brief Row types as the reference spells them (core.hpp:1050-1065): the
/// bare type for one variable, a tuple otherwise.
template <Kernel K>
using input_type = typename detail::maybe_single<input_tuple<K>>::type;
template <Kernel K>
using output_type = typename detail::maybe_single<output_tuple<K>>::type;

/// @brief Host columns of a kernel's inputs or outputs (structure of arrays).
template <typename Tuple>
struct Columns;

template <typename... Ts>
struct Columns<std::tuple<Ts...>>
{
    std::tuple<std::vector<Ts>...> cols;

    std::size_t
    size() const
    {
        return std::get<0>(cols).size();
    }

    void
    resize(std::size_t n)
    {
        std::apply([n](auto&... c) { (c.resize(n), ...); }, cols);
    }

    std::vector<const void*>
    cptrs() const
    {
        return std::apply([](const auto&... c)
                          { return std::vector<const void*>{static_cast<const void*>(c.data())...}; },
                          cols);
    }

    std::vector<void*>
    ptrs()
    {
        return std::apply([](auto&... c)
                          { return std::vector<void*>{static_cast<void*>(c.data())...}; }, cols);
    }

    std::tuple<Ts...>
    row(std::size_t i) const
    {
        return std::apply([i](const auto&... c) { return std::tuple<Ts...>{c[i]...}; }, cols);
    }

    void
    set_row(std::size_t i, const std::tuple<Ts...>& v)
    {
        [&]<std::size_t... I>(std::index_sequence<I...>)
        { ((std::get<I>(cols)[i] = std::get<I>(v)), ...); }(std::index_sequence_for<Ts...>{});
    }

    /// @brief Row in the reference spelling (bare value for one column).
    auto
    row_value(std::size_t i) const
    {
        if constexpr(sizeof...(Ts) == 1)
            return std::get<0>(cols)[i];
        else
            return row(i);
    }

    template <typename V>
    void
    set_row_value(std::size_t i, const V& v)
    {
        if constexpr(sizeof...(Ts) == 1 && !requires { std::tuple_size<V>::value; })
            std::get<0>(cols)[i] = static_cast<std::tuple_element_t<0, std::tuple<Ts...>>>(v);
        else
            set_row(i, std::tuple<Ts...>(v));
    }
};

/// @brief Function-like invocation (reference compact_function /
/// hybrid_function): every thread runs to HALT; outputs keep input order.
template <Kernel K, int Mode>
class function
{
    using in_t  = input_tuple<K>;
    using out_t = output_tuple<K>;

    std::size_t      m_n = 0;
    params           m_p;
    Columns<in_t>    m_in;
    Columns<out_t>   m_out;
    loom_run_stats   m_stats{};

    template <std::size_t I, typename Arg>
    void
    set_one(Arg&& a)
    {
        using T = std::tuple_element_t<I, in_t>;
        auto& c = std::get<I>(m_in.cols);
        if constexpr(detail::Column<Arg>)
        {
            c.assign(std::ranges::begin(a), std::ranges::end(a));
        }
        else
            c.assign(m_n, static_cast<T>(a));
    }

  public:
    using program = K;

    explicit function(std::size_t n = 0, params p = {}) noexcept : m_n(n), m_p(p) {}

    /// @brief Columns (ranges) or scalars (broadcast), one per input
    /// (reference MemoryFacade::set_inputs, bve.hpp:1125-1133).
    template <typename... Args>
        requires(sizeof...(Args) == std::tuple_size_v<in_t>)
    void
    set_inputs(Args&&... args) noexcept
    {
        // size: the first column argument, else the constructor size
        std::size_t n = m_n;
        (
            [&]
            {
                if constexpr(detail::Column<Args>)
                    if(n == m_n) n = static_cast<std::size_t>(std::ranges::size(args));
            }(),
            ...);
        m_n = n;
        [&]<std::size_t... I>(std::index_sequence<I...>)
        { (set_one<I>(std::forward<Args>(args)), ...); }(std::index_sequence_for<Args...>{});
    }

    /// @brief Run all threads to HALT on the GPU.
    void
    run() noexcept
    {
        m_out.resize(m_n);
        const loom_params c = detail::to_c(m_p, Mode);
        auto              in = m_in.cptrs();
        auto              out = m_out.ptrs();
        detail::check(loom_run_host(kernel_index<K>(), Mode, in.data(), out.data(), m_n, &c,
                                    &m_stats),
                      "loom_run_host");
    }

    /// @brief `f(args...)`: set inputs, run, return outputs (a tuple of
    /// columns, or rows with zip_outputs; bve.hpp:1135-1173).
    template <bool zip_outputs = false, typename... Args>
    auto
    operator()(Args&&... args) noexcept
    {
        set_inputs(std::forward<Args>(args)...);
        run();
        if constexpr(zip_outputs)
            return get_output_zipped();
        else
            return get_outputs();
    }

    const auto&
    get_outputs() const noexcept
    {
        return m_out.cols;
    }

    auto
    get_output_zipped() const noexcept
    {
        std::vector<output_type<K>> rows(m_out.size());
        for(std::size_t i = 0; i < rows.size(); ++i) rows[i] = m_out.row_value(i);
        return rows;
    }

    void
    reset() noexcept
    {
        m_out.resize(0);
    }

    bool
    done() const noexcept
    {
        return m_out.size() == m_n;
    }

    std::size_t
    size() const noexcept
    {
        return m_n;
    }

    const loom_run_stats&
    stats() const noexcept
    {
        return m_stats;
    }
};

template <Kernel K>
using compact_function = function<K, LOOM_MODE_COMPACT>;

template <Kernel K>
using hybrid_function = function<K, LOOM_MODE_HYBRID>;

/// @brief Streaming invocation (reference hybrid_stream with the
/// concurrent stream facade protocol, simd.hpp:2797-2903,
/// bve.hpp:1208-1229), over the device-resident stream of the C ABI
/// (loom_stream_*): `slots` thread slots whose registers stay on the GPU;
/// `set_inputs` fills them, `s()` advances the live slots (each wave until
/// at most `live_threshold` of its lanes are live), `sort()` reports
/// whether halted slots can be exchanged, `exchange(rest)` swaps the outputs
/// of halted threads for new inputs (one for one, only those rows cross
/// PCIe), `done()` is true when no thread is live, `get_output_zipped()`
/// returns the outputs still held.  Outputs come in the order threads
/// halted, as the reference's come in its sorted order.  Calls may come from
/// two threads at once (a worker stepping, the main thread exchanging, as
/// mp::Context does, mp.hpp:67-131): the C ABI serialises them.
template <Kernel K>
class hybrid_stream
{
    using in_t  = input_tuple<K>;
    using out_t = output_tuple<K>;
    using row_t = output_type<K>;

    loom_stream*               m_s = nullptr;
    std::size_t                m_slots;
    Columns<in_t>              m_in;  // staging of rows going in
    Columns<out_t>             m_out; // staging of rows coming out
    std::vector<std::uint64_t> m_ids;
    loom_run_stats             m_stats{};

    template <typename It, typename End>
    std::size_t
    stage(It& it, const End& end, std::size_t cap) noexcept
    {
        std::size_t m = 0;
        for(; m < cap && it != end; ++m, ++it) m_in.set_row_value(m, *it);
        return m;
    }

    std::vector<row_t>
    rows(std::size_t m) const noexcept
    {
        std::vector<row_t> res(m);
        for(std::size_t i = 0; i < m; ++i) res[i] = m_out.row_value(i);
        return res;
    }

  public:
    using program = K;

    explicit hybrid_stream(std::size_t slots, params p = {}) noexcept :
        m_slots((std::max<std::size_t>(slots, 1) + 63) / 64 * 64)
    {
        const loom_params c = detail::to_c(p, LOOM_MODE_HYBRID);
        detail::check(loom_stream_create(kernel_index<K>(), m_slots, &c, &m_s),
                      "loom_stream_create");
        m_in.resize(m_slots);
        m_out.resize(m_slots);
        m_ids.resize(m_slots);
    }

    hybrid_stream(const hybrid_stream&)            = delete;
    hybrid_stream& operator=(const hybrid_stream&) = delete;

    ~hybrid_stream() { loom_stream_destroy(m_s); }

    /// @brief Start the first rows of @p input in the slots (the stream
    /// starts over); returns the iterator past the rows taken.
    template <std::ranges::input_range R>
    auto
    set_inputs(R&& input) noexcept
    {
        auto              it = std::ranges::begin(input);
        const std::size_t m  = stage(it, std::ranges::end(input), m_slots);
        auto              ip = m_in.cptrs();
        std::uint64_t     taken = 0;
        detail::check(loom_stream_set_inputs(m_s, ip.data(), m, &taken), "loom_stream_set_inputs");
        return it;
    }

    /// @brief One step of every live thread on the GPU.
    void
    operator()() noexcept
    {
        detail::check(loom_stream_step(m_s, &m_stats), "loom_stream_step");
    }

    /// @brief True when halted threads' outputs can be exchanged.
    bool
    sort() noexcept
    {
        int avail = 0;
        detail::check(loom_stream_sort(m_s, &avail), "loom_stream_sort");
        return avail != 0;
    }

    /// @brief Outputs of halted threads, their slots given the next rows of
    /// @p input; as many as both allow.
    template <std::ranges::input_range R>
    std::vector<row_t>
    exchange(R&& input) noexcept
    {
        std::uint64_t live = 0, halted = 0;
        detail::check(loom_stream_state(m_s, &live, &halted), "loom_stream_state");
        auto              it = std::ranges::begin(input);
        const std::size_t m  = stage(it, std::ranges::end(input), halted);
        if(m == 0) return {};
        auto          ip = m_in.cptrs();
        auto          op = m_out.ptrs();
        std::uint64_t got = 0;
        detail::check(loom_stream_exchange(m_s, ip.data(), m, op.data(), m_ids.data(), &got),
                      "loom_stream_exchange");
        return rows(got);
    }

    bool
    done() const noexcept
    {
        std::uint64_t live = 0;
        detail::check(loom_stream_state(m_s, &live, nullptr), "loom_stream_state");
        return live == 0;
    }

    /// @brief Outputs still held by halted threads (their slots become empty).
    std::vector<row_t>
    get_output_zipped() noexcept
    {
        std::uint64_t halted = 0;
        detail::check(loom_stream_state(m_s, nullptr, &halted), "loom_stream_state");
        auto          op  = m_out.ptrs();
        std::uint64_t got = 0;
        detail::check(loom_stream_drain(m_s, op.data(), m_ids.data(), halted, &got),
                      "loom_stream_drain");
        return rows(got);
    }

    std::size_t
    size() const noexcept
    {
        return m_slots;
    }

    const loom_run_stats&
    stats() const noexcept
    {
        return m_stats;
    }

    loom_stream_counters
    counters() const noexcept
    {
        loom_stream_counters c{};
        detail::check(loom_stream_counters_get(m_s, &c), "loom_stream_counters_get");
        return c;
    }
};

/// @brief Kernel execution context of a worker (reference mp::Context,
/// mp.hpp:61-131): a stream of one kernel, a lock that orders sorting
/// against exchanging, and the flag that an exchange is available.  A
/// worker thread steps and sorts (`try_step`), the main thread exchanges
/// (`try_exchange`); stepping may run while an exchange does.
template <template <typename> class Stream, typename K>
struct basic_context
{
    Stream<K>        stream;
    std::mutex       lock;
    bool             need_sort = false;
    std::atomic_flag exchange_available;

    template <typename... A>
    explicit basic_context(A&&... a) noexcept : stream(std::forward<A>(a)...)
    {
    }

    void
    try_sort() noexcept
    {
        if(need_sort && lock.try_lock())
        {
            const bool ex = stream.sort();
            lock.unlock();
            need_sort = false;
            if(ex) exchange_available.test_and_set(std::memory_order_relaxed);
        }
    }

    void
    try_step() noexcept
    {
        if(!(need_sort || stream.done()))
        {
            stream();
            need_sort = true;
        }
        try_sort();
    }

    template <typename It, typename Sentinel, typename Out>
    bool
    try_exchange(It& ib, const Sentinel s, Out& out) noexcept
    {
        if(ib != s && exchange_available.test(std::memory_order_relaxed) && lock.try_lock())
        {
            for(const auto& ex: stream.exchange(std::ranges::subrange(ib, s)))
            {
                *out = ex;
                ++out;
                ++ib;
            }
            lock.unlock();
            exchange_available.clear(std::memory_order_relaxed);
            return true;
        }
        return false;
    }
};

/// @brief Several host threads feeding one GPU (reference mp::Shop,
/// mp.hpp:148-385, with a dynamic number of workers): every worker owns one
/// context (a device-resident stream) per kernel and round-robins over them,
/// stepping and sorting; the main thread binds the inputs, exchanges until
/// every input is taken, signals the workers (`set()`), which finish their
/// streams, and collects the last outputs (`finalize`).  Each stream has a
/// HIP stream of its own, so the workers' launches overlap on the device.
template <template <typename> class Stream, typename... Ks>
class basic_shop
{
    using bench_t = std::tuple<std::unique_ptr<basic_context<Stream, Ks>>...>;
    std::vector<bench_t> m_benches;
    std::atomic<bool>    m_signal{false};

    template <typename F>
    void
    visit(F&& f)
    {
        for(auto& b: m_benches) f(b);
    }

  public:
    /// @brief Joins the worker threads when it goes out of scope (the
    /// reference thread_pool).
    class thread_pool
    {
        std::vector<std::thread> m_threads;

      public:
        explicit thread_pool(std::vector<std::thread>&& t) noexcept : m_threads(std::move(t)) {}
        thread_pool(thread_pool&&) noexcept = default;
        ~thread_pool()
        {
            for(auto& t: m_threads)
                if(t.joinable()) t.join();
        }
        std::size_t
        size() const noexcept
        {
            return m_threads.size();
        }
    };

    template <typename... A>
    basic_shop(std::size_t workers, A&&... stream_args) noexcept
    {
        m_benches.resize(std::max<std::size_t>(workers, 1));
        for(auto& b: m_benches)
            std::apply([&](auto&... ctx)
                       { ((ctx = std::make_unique<typename std::remove_reference_t<decltype(ctx)>::element_type>(
                               stream_args...)),
                          ...); },
                       b);
    }

    /// @brief Signal the workers that no input is left (reference Signal::set).
    void
    set() noexcept
    {
        m_signal.store(true, std::memory_order_release);
    }

    /// @brief Start one thread per bench; each steps its contexts until the
    /// signal is set and every stream of its bench is done.
    [[nodiscard]] thread_pool
    get_thread_pool()
    {
        m_signal.store(false, std::memory_order_release);
        std::vector<std::thread> t;
        for(auto& b: m_benches)
            t.emplace_back(
                [this, &b]
                {
                    std::array<bool, sizeof...(Ks)> done{};
                    while(!std::all_of(done.begin(), done.end(), [](bool d) { return d; }))
                        [&]<std::size_t... I>(std::index_sequence<I...>)
                        {
                            (
                                [&]
                                {
                                    auto& ctx = *std::get<I>(b);
                                    if(done[I]) return;
                                    if(m_signal.load(std::memory_order_acquire) && ctx.lock.try_lock())
                                    {
                                        if(ctx.stream.done()) done[I] = true;
                                        ctx.lock.unlock();
                                    }
                                    ctx.try_step();
                                }(),
                                ...);
                        }(std::index_sequence_for<Ks...>{});
                });
        return thread_pool(std::move(t));
    }

    /// @brief Give every stream its first inputs; returns, per kernel, the
    /// pair (next input, end) of what is left.
    template <typename... Inputs>
        requires(sizeof...(Inputs) == sizeof...(Ks))
    [[nodiscard]] auto
    bind_inputs(Inputs&... inputs)
    {
        auto ret = std::make_tuple(std::make_pair(std::ranges::begin(inputs), std::ranges::end(inputs))...);
        visit(
            [&](auto& bench)
            {
                [&]<std::size_t... I>(std::index_sequence<I...>)
                {
                    ((std::get<I>(ret).first = std::get<I>(bench)->stream.set_inputs(
                          std::ranges::subrange(std::get<I>(ret).first, std::get<I>(ret).second))),
                     ...);
                }(std::index_sequence_for<Ks...>{});
            });
        return ret;
    }

    /// @brief One exchange attempt on every context; true if any exchanged.
    template <typename InputCollection, typename OutputCollection>
    bool
    exchange(InputCollection& ic, OutputCollection& oc) noexcept
    {
        bool ret = false;
        visit(
            [&](auto& bench)
            {
                [&]<std::size_t... I>(std::index_sequence<I...>)
                {
                    ((ret = std::get<I>(bench)->try_exchange(std::get<I>(ic).first,
                                                             std::get<I>(ic).second,
                                                             std::get<I>(oc)) ||
                            ret),
                     ...);
                }(std::index_sequence_for<Ks...>{});
            });
        return ret;
    }

    /// @brief The outputs still held by every stream.
    template <typename OutputCollection>
    void
    finalize(OutputCollection& oc)
    {
        visit(
            [&](auto& bench)
            {
                [&]<std::size_t... I>(std::index_sequence<I...>)
                {
                    (std::ranges::copy(std::get<I>(bench)->stream.get_output_zipped(),
                                       std::get<I>(oc)),
                     ...);
                }(std::index_sequence_for<Ks...>{});
            });
    }

    std::size_t
    workers() const noexcept
    {
        return m_benches.size();
    }
};

/// @brief Input collection complete: every kernel's inputs were taken
/// (reference mp::is_complete).
template <typename InputCollection>
bool
is_complete(const InputCollection& ic) noexcept
{
    return std::apply([](const auto&... p) { return (... && (p.first == p.second)); }, ic);
}

template <Kernel K>
using context = basic_context<hybrid_stream, K>;

template <Kernel... Ks>
using shop = basic_shop<hybrid_stream, Ks...>;

} // namespace loom::vm::hip
