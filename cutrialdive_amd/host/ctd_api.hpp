// ctd_api.hpp -- C++ host interface of the MI355X trial-factoring engine.
//
// Keeps cutrialdive's host-side API so that callers of the reference compile
// against it unchanged:
//   * number_sequence concepts          (number_sequence.hpp:19-58, number_sequence_detail.hpp:44-274)
//   * num_seq_id / num_seq_spec         (num_seq_id.hpp:16-27, num_seq_spec.hpp:16-22)
//   * dispatch_num_seq                  (num_seq_dispatch.hpp:119-168)
//   * trial_factoring_options           (trial_factoring_options.hpp:25-39)
//   * factors_buffer / factoring_results (factors_buffer.hpp:33-96, factoring_results.hpp:146-201)
//   * trial_factor / resume_trial_factoring (trial_factoring.hpp:20-36)
// The compute goes through the C ABI of include/ctd_hip.h: the built-in
// sequence families name the device kernel that propagates them with
// `device_seq_id` and `device_param()`.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <concepts>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <optional>
#include <ostream>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include "ctd_hip.h"

namespace cutrialdive {

// ---------------------------------------------------------------- verbosity / logging

enum class verbosity : uint16_t { silent = 0, normal = 1, verbose = 2, debug = 3, diagnostic = 4 };
inline bool is_valid_verbosity(uint16_t v) { return v <= 4; }

struct logger {
    verbosity level = verbosity::normal;
    std::ostream *out = nullptr;
    bool on(verbosity v) const { return out && level >= v; }
};
logger &log();

// ---------------------------------------------------------------- sequence ids and specs

enum class num_seq_id : uint32_t {
    mersenne,
    proth,
    riesel,
    smarandache,
    fibonacci,
    lucas,
    proth_fixed_n,
    riesel_fixed_n,
    tribonacci
};

num_seq_id num_seq_id_from_string(std::string const &s);
std::string to_string(num_seq_id id);

struct num_seq_spec {
    num_seq_id seq_id;
    std::string seq_params;
    bool operator==(num_seq_spec const &) const = default;
};

// ---------------------------------------------------------------- plugin concepts

struct no_state_t {};
struct no_mu_t {};
struct mu_both_t {
    unsigned __int128 mu128;
    uint64_t mu64;
};

template <typename T>
decltype(auto) get_math_sequence(T &&seq)
{
    if constexpr (requires { seq.math_sequence(); })
        return seq.math_sequence();
    else
        return std::forward<T>(seq);
}

namespace details {
template <typename Seq>
using math_sequence_type = std::remove_cvref_t<decltype(get_math_sequence(std::declval<Seq>()))>;
}

template <typename Seq> concept HasState = Seq::has_state;
template <typename Seq> concept InitializeFromValue = Seq::initialize_from_value;
template <typename Seq>
concept HasValue = requires(Seq s, typename Seq::index_type n) {
    { s.value(n) } -> std::same_as<typename Seq::value_type>;
};
template <typename Seq>
concept HasValueMod = requires(Seq s, typename Seq::index_type n, typename Seq::residue_type d) {
    { s.value_mod(n, d) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasValueModWithState =
    requires(Seq s, typename Seq::index_type n, typename Seq::residue_type d, typename Seq::state_type &st) {
        { s.value_mod(n, d, st) } -> std::same_as<typename Seq::residue_type>;
    };
template <typename Seq>
concept HasValueModMu = requires(Seq s, typename Seq::index_type n, typename Seq::residue_type d, typename Seq::mu_type mu) {
    { s.value_mod_mu(n, d, mu) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasValueModMuWithState = requires(Seq s, typename Seq::index_type n, typename Seq::residue_type d,
                                          typename Seq::mu_type mu, typename Seq::state_type &st) {
    { s.value_mod_mu(n, d, mu, st) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueMod = requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n,
                                   typename Seq::residue_type d) {
    { s.next_value_mod(r, n, d) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueModWithState = requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n,
                                            typename Seq::residue_type d, typename Seq::state_type &st) {
    { s.next_value_mod(r, n, d, st) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueModMu = requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n,
                                     typename Seq::residue_type d, typename Seq::mu_type mu) {
    { s.next_value_mod_mu(r, n, d, mu) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueModMuWithState =
    requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n, typename Seq::residue_type d,
             typename Seq::mu_type mu, typename Seq::state_type &st) {
        { s.next_value_mod_mu(r, n, d, mu, st) } -> std::same_as<typename Seq::residue_type>;
    };
template <typename Seq>
concept HasValueModTwo = requires(Seq s, typename Seq::index_type n) {
    { s.value_mod_2(n) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasValueModTwoWithState = requires(Seq s, typename Seq::index_type n, typename Seq::state_type &st) {
    { s.value_mod_2(n, st) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueModTwo = requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n) {
    { s.next_value_mod_2(r, n) } -> std::same_as<typename Seq::residue_type>;
};
template <typename Seq>
concept HasNextValueModTwoWithState =
    requires(Seq s, typename Seq::residue_type r, typename Seq::index_type n, typename Seq::state_type &st) {
        { s.next_value_mod_2(r, n, st) } -> std::same_as<typename Seq::residue_type>;
    };
template <typename Seq>
concept HasShortName = requires(Seq s) {
    { s.short_name() } -> std::convertible_to<std::string_view>;
};
template <typename Seq>
concept HasSequenceShortName = HasShortName<Seq> || HasShortName<details::math_sequence_type<Seq>>;
template <typename Seq>
concept HasExpressionPrinter = requires(Seq s, typename details::math_sequence_type<Seq>::index_type n, std::ostream &o) {
    { get_math_sequence(s).print_expression(n, o) } -> std::same_as<std::ostream &>;
};

// The same requirement set as number_sequence.hpp:19-53.
template <typename Seq>
concept PureMathSequence =
    requires { typename Seq::index_type; typename Seq::residue_type; } &&
    ((InitializeFromValue<Seq> && HasValue<Seq>) || HasValueMod<Seq> || HasValueModMu<Seq> ||
     (HasState<Seq> && (HasValueModWithState<Seq> || HasValueModMuWithState<Seq>))) &&
    ((InitializeFromValue<Seq> && HasValue<Seq>) || HasValueMod<Seq> || HasValueModTwo<Seq> ||
     (HasState<Seq> && (HasValueModWithState<Seq> || HasValueModTwoWithState<Seq>))) &&
    (HasNextValueMod<Seq> || HasNextValueModMu<Seq> ||
     (HasState<Seq> && (HasNextValueModWithState<Seq> || HasNextValueModMuWithState<Seq>))) &&
    (HasNextValueMod<Seq> || HasNextValueModTwo<Seq> ||
     (HasState<Seq> && (HasNextValueModWithState<Seq> || HasNextValueModTwoWithState<Seq>)));

template <typename Seq>
concept NumberSequence = PureMathSequence<details::math_sequence_type<Seq>> && HasSequenceShortName<Seq>;

// A pure math sequence the device engine propagates natively: it names its
// kernel (a CTD_SEQ_* id) and the kernel's numeric parameter.
template <typename Seq>
concept DeviceSequence = PureMathSequence<Seq> && requires(Seq s) {
    { Seq::device_seq_id } -> std::convertible_to<uint32_t>;
    { s.device_param() } -> std::convertible_to<uint64_t>;
};

// ---------------------------------------------------------------- options and results

struct runtime_options {
    uint32_t threads_count = 4;
    verbosity verbosity_level = verbosity::normal;
    static runtime_options default_options() { return {}; }
};

struct trial_factoring_options {
    uint64_t n0{};
    uint64_t n1{};
    uint64_t f0{};
    uint64_t f1{};
    std::optional<std::filesystem::path> output_path;
    uint32_t max_factors_per_number{32};
    bool operator==(trial_factoring_options const &) const = default;
};
std::ostream &operator<<(std::ostream &out, trial_factoring_options const &o);

struct tf_runtime_options {
    runtime_options common_options;
    std::chrono::seconds checkpoint_period{360};
    std::chrono::milliseconds progress_period{1000};
    bool is_progress_enabled = true;
    uint64_t segment_length = uint64_t{1} << 26;
    int32_t device_id = 0;
    std::optional<int32_t> grid_size;
    std::optional<int32_t> block_size;
    static tf_runtime_options default_options() { return {}; }
};

template <typename PrimeT, typename ExponentT>
struct factor {
    PrimeT prime;
    ExponentT exponent;
    factor(PrimeT p, ExponentT e) : prime(p), exponent(e) {}
};

template <typename PrimeT, typename ExponentT>
std::ostream &operator<<(std::ostream &out, factor<PrimeT, ExponentT> const &f)
{
    out << f.prime;
    if (f.exponent > ExponentT{1}) out << "^" << f.exponent;
    return out;
}

// factoring_results (factoring_results.hpp:146-201)
template <typename ValueT, typename ExponentT>
class factoring_results {
public:
    factoring_results(uint64_t n0, std::size_t count) : n0_(n0) { locations_.reserve(count); }
    uint64_t n0() const { return n0_; }
    std::size_t size() const { return locations_.size(); }
    bool empty() const { return locations_.empty(); }
    void reserve(std::size_t n) { factors_.reserve(n); }
    template <typename PrimeT>
    void add_factors(std::span<PrimeT> fs, uint32_t found)
    {
        std::size_t off = factors_.size();
        for (auto p : fs) factors_.emplace_back(p, ExponentT{1});
        locations_.push_back({off, fs.size()});
        max_factors_count_ = std::max(max_factors_count_, found);
        factors_count_ += found;
    }
    std::span<factor<ValueT, ExponentT> const> operator[](std::size_t i) const
    {
        return {factors_.data() + locations_[i].offset, locations_[i].length};
    }
    uint32_t max_factors_count() const { return max_factors_count_; }
    uint64_t factors_count() const { return factors_count_; }

private:
    struct location {
        std::size_t offset, length;
    };
    uint64_t n0_;
    std::vector<location> locations_;
    std::vector<factor<ValueT, ExponentT>> factors_;
    uint32_t max_factors_count_{};
    uint64_t factors_count_{};
};

// "<index> - p1, p2\n" per index (factoring_results.hpp:183-201)
template <typename V, typename E>
std::ostream &operator<<(std::ostream &out, factoring_results<V, E> const &r)
{
    for (std::size_t i = 0; i < r.size(); ++i) {
        out << r.n0() + i << " - ";
        const char *sep = "";
        for (auto const &f : r[i]) {
            out << sep << f;
            sep = ", ";
        }
        out << "\n";
    }
    return out;
}

// factors_buffer: counts[N] and factors[N * max] index-major (factors_buffer.hpp:77-86)
template <typename PrimeT>
class factors_buffer {
public:
    factors_buffer(uint64_t n0, uint64_t count, uint32_t max)
        : counts_(count), factors_(count * max), max_(max), count_(count), n0_(n0)
    {
    }
    factors_buffer(factors_buffer &&) noexcept = default;
    factors_buffer &operator=(factors_buffer &&) noexcept = default;
    factors_buffer(factors_buffer const &) = delete;
    factors_buffer &operator=(factors_buffer const &) = delete;

    void push_factor(uint64_t idx, PrimeT p)
    {
        uint32_t pos = std::atomic_ref<uint32_t>{counts_[idx]}.fetch_add(1, std::memory_order_relaxed);
        if (pos < max_) factors_[idx * max_ + pos] = p;
    }
    template <typename ResultPrimeT, typename ExponentT>
    factoring_results<ResultPrimeT, ExponentT> to_factoring_results() const
    {
        factoring_results<ResultPrimeT, ExponentT> r{n0_, count_};
        for (std::size_t i = 0; i < count_; ++i) {
            uint32_t k = std::min(counts_[i], max_);
            std::span<PrimeT> s{&factors_[i * max_], k};
            std::sort(s.begin(), s.end());
            r.add_factors(s, counts_[i]);
        }
        return r;
    }
    uint64_t n0() const { return n0_; }
    uint64_t numbers_count() const { return count_; }
    uint32_t max_factors_per_number() const { return max_; }
    void report_excess_factors_if_any(std::ostream &out);

    std::vector<uint32_t> &counts() { return counts_; }
    std::vector<PrimeT> &factors() { return factors_; }
    std::vector<uint32_t> const &counts() const { return counts_; }
    std::vector<PrimeT> const &factors() const { return factors_; }

private:
    std::vector<uint32_t> counts_;
    mutable std::vector<PrimeT> factors_;
    uint32_t max_;
    uint64_t count_;
    uint64_t n0_;
    std::optional<uint32_t> last_max_;
};

// ---------------------------------------------------------------- checkpointing (checkpoint.hpp:22-110)

class factors_buffer_holder {
public:
    explicit factors_buffer_holder(factors_buffer<uint64_t> *buf) : ptr_(buf) {}
    explicit factors_buffer_holder(factors_buffer<uint64_t> buf) : owned_(std::move(buf)) { ptr_ = &*owned_; }
    factors_buffer_holder(factors_buffer_holder &&o) noexcept : owned_(std::move(o.owned_))
    {
        ptr_ = owned_ ? &*owned_ : o.ptr_;
    }
    factors_buffer<uint64_t> &get() { return *ptr_; }
    factors_buffer<uint64_t> const &get() const { return *ptr_; }

private:
    factors_buffer<uint64_t> *ptr_ = nullptr;
    std::optional<factors_buffer<uint64_t>> owned_;
};

struct job_spec {
    num_seq_spec seq_spec;
    trial_factoring_options tf_options;
};

struct engine_state {
    uint64_t last_processed_prime;
    factors_buffer_holder factors_buffer;
};

struct checkpoint_data {
    job_spec spec;
    engine_state state;
};

class checkpoint_manager {
public:
    checkpoint_manager(std::filesystem::path path, std::chrono::seconds period, job_spec spec);
    std::filesystem::path checkpoint_path() const { return path_; }
    std::chrono::seconds period() const { return period_; }
    bool due() const;
    void write(engine_state const &state);
    static std::optional<checkpoint_data> load(std::filesystem::path const &path);
    void remove_checkpoint();

private:
    std::filesystem::path path_;
    std::chrono::seconds period_;
    job_spec spec_;
    std::chrono::steady_clock::time_point last_write_;
};

// ---------------------------------------------------------------- the engine seam

struct trial_factoring_context {
    trial_factoring_options const &options;
    tf_runtime_options const &runtime_options;
    factoring_results<uint64_t, uint32_t> &results;
    std::ostream &output_stream;
    checkpoint_manager *checkpoint{};
    engine_state *resume_state{};
};

// device_trial_factor (trial_factoring.cuh:336-513) for any device sequence:
// runs the job through ctd_hip_trial_factor.
void device_trial_factor_raw(trial_factoring_context &ctx, uint32_t seq_id, uint64_t param);

template <typename Seq>
    requires DeviceSequence<Seq>
void device_trial_factor(trial_factoring_context &ctx, Seq seq)
{
    device_trial_factor_raw(ctx, Seq::device_seq_id, seq.device_param());
}

void trial_factor_front_begin(trial_factoring_context &ctx, bool &proceed, std::unique_ptr<std::ostream> &file);
void trial_factor_front_end(trial_factoring_context &ctx, std::unique_ptr<std::ostream> &file);

// trial_factor<Seq> (trial_factoring_impl.hpp:369-441)
template <typename NumberSequenceT>
    requires NumberSequence<NumberSequenceT>
void trial_factor(trial_factoring_context &&ctx, auto &&...args)
{
    bool proceed = false;
    std::unique_ptr<std::ostream> file;
    trial_factor_front_begin(ctx, proceed, file);
    if (!proceed) return;
    {
        auto t0 = std::chrono::steady_clock::now();
        NumberSequenceT seq{std::forward<decltype(args)>(args)...};
        auto pure = get_math_sequence(seq);
        static_assert(DeviceSequence<std::remove_cvref_t<decltype(pure)>>,
                      "this engine propagates sequences with a device kernel (device_seq_id / device_param)");
        device_trial_factor(ctx, pure);
        ctx.output_stream << "Trial factoring took "
                          << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << "s"
                          << std::endl;
    }
    trial_factor_front_end(ctx, file);
}

factoring_results<uint64_t, uint32_t> trial_factor(num_seq_spec spec, trial_factoring_options const &opts,
                                                  tf_runtime_options const &rt, std::ostream &out);
factoring_results<uint64_t, uint32_t> resume_trial_factoring(std::filesystem::path const &checkpoint_path,
                                                             tf_runtime_options const &rt, std::ostream &out);

} // namespace cutrialdive

#include "ctd_sequences.hpp"
