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
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <ranges>
#include <tuple>
#include <type_traits>
#include <vector>

namespace loom::vm::hip
{
/// @brief Backend parameters (see loom_params).
struct params
{
    std::uint32_t threshold      = 0;
    std::uint32_t live_threshold = 48;
    bool          refill         = true;
    bool          regroup        = false;
    int           reconv         = -1; ///< -1: per mode default
    std::uint32_t waves_per_cu   = 0;
    std::uint64_t queue_capacity = 0;
    std::uint64_t max_steps      = 0;
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
    c.reconv         = p.reconv >= 0 ? static_cast<std::uint32_t>(p.reconv) :
                                       (mode == LOOM_MODE_HYBRID ? 2u : 0u);
    c.waves_per_cu   = p.waves_per_cu;
    c.queue_capacity = p.queue_capacity;
    c.max_steps      = p.max_steps;
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
/// ConcurrentStreamFacade protocol, simd.hpp:2797-2903, bve.hpp:1208-1229):
/// a pool of `slots` resident threads; `set_inputs` fills it, `s()` runs the
/// admitted threads on the GPU, `sort()` reports whether halted slots can be
/// exchanged, `exchange(rest)` swaps the outputs of halted slots for new
/// inputs (one for one), `done()` is true when no admitted thread is left.
template <Kernel K>
class hybrid_stream
{
    using in_t  = input_tuple<K>;
    using out_t = output_tuple<K>;
    using row_t = output_type<K>;

    std::size_t       m_slots;
    params            m_p;
    Columns<in_t>     m_in;     // slot inputs
    Columns<out_t>    m_out;    // slot outputs
    std::vector<char> m_state;  // 0 empty, 1 admitted, 2 halted (holds outputs)

  public:
    using program = K;

    explicit hybrid_stream(std::size_t slots, params p = {}) noexcept : m_slots(slots), m_p(p)
    {
        m_in.resize(slots);
        m_out.resize(slots);
        m_state.assign(slots, 0);
    }

    /// @brief Fill empty slots from @p input (rows of the input type);
    /// returns the iterator past the consumed rows.
    template <std::ranges::input_range R>
    auto
    set_inputs(R&& input) noexcept
    {
        auto it = std::ranges::begin(input);
        for(std::size_t s = 0; s < m_slots && it != std::ranges::end(input); ++s)
            if(m_state[s] == 0)
            {
                m_in.set_row_value(s, *it);
                m_state[s] = 1;
                ++it;
            }
        return it;
    }

    /// @brief Run every admitted thread to HALT (one GPU launch).
    void
    operator()() noexcept
    {
        std::vector<std::size_t> idx;
        for(std::size_t s = 0; s < m_slots; ++s)
            if(m_state[s] == 1) idx.push_back(s);
        if(idx.empty()) return;
        Columns<in_t>  in;
        Columns<out_t> out;
        in.resize(idx.size());
        out.resize(idx.size());
        for(std::size_t i = 0; i < idx.size(); ++i) in.set_row(i, m_in.row(idx[i]));
        const loom_params c    = detail::to_c(m_p, LOOM_MODE_HYBRID);
        auto              ip   = in.cptrs();
        auto              op   = out.ptrs();
        loom_run_stats    st{};
        detail::check(loom_run_host(kernel_index<K>(), LOOM_MODE_HYBRID, ip.data(), op.data(),
                                    idx.size(), &c, &st),
                      "loom_run_host");
        for(std::size_t i = 0; i < idx.size(); ++i)
        {
            [&]<std::size_t... I>(std::index_sequence<I...>)
            { ((std::get<I>(m_out.cols)[idx[i]] = std::get<I>(out.cols)[i]), ...); }(
                std::make_index_sequence<std::tuple_size_v<out_t>>{});
            m_state[idx[i]] = 2;
        }
    }

    /// @brief True when halted slots hold outputs to exchange.
    bool
    sort() noexcept
    {
        return std::ranges::count(m_state, char(2)) > 0;
    }

    /// @brief Outputs of halted slots, each slot refilled with the next row
    /// of @p input; as many as both allow.
    template <std::ranges::input_range R>
    std::vector<row_t>
    exchange(R&& input) noexcept
    {
        std::vector<row_t> res;
        auto               it = std::ranges::begin(input);
        for(std::size_t s = 0; s < m_slots && it != std::ranges::end(input); ++s)
            if(m_state[s] == 2)
            {
                res.push_back(m_out.row_value(s));
                m_in.set_row_value(s, *it);
                m_state[s] = 1;
                ++it;
            }
        return res;
    }

    bool
    done() const noexcept
    {
        return std::ranges::count(m_state, char(1)) == 0;
    }

    /// @brief Outputs still held by halted slots (then the slots are empty).
    std::vector<row_t>
    get_output_zipped() noexcept
    {
        std::vector<row_t> res;
        for(std::size_t s = 0; s < m_slots; ++s)
            if(m_state[s] == 2)
            {
                res.push_back(m_out.row_value(s));
                m_state[s] = 0;
            }
        return res;
    }

    std::size_t
    size() const noexcept
    {
        return m_slots;
    }
};
} // namespace loom::vm::hip
