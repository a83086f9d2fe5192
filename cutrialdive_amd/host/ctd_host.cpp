// ctd_host.cpp -- host side of the trial-factoring engine: the public API
// (trial_factoring.cpp:18-92), the engine front (trial_factoring_impl.hpp:369-441),
// the device seam (trial_factoring.cuh:336-513) over the C ABI, and the
// checkpoint v1 file format (checkpoint.cpp:21-231).
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <unordered_map>

#include "ctd_api.hpp"
#include "ctd_progress.hpp"

namespace fs = std::filesystem;

namespace cutrialdive {

logger &log()
{
    static logger l{verbosity::normal, &std::cout};
    return l;
}

// num_seq_id_from_string / to_string (num_seq_id.cpp:25-43, num_seq_id.hpp:42-57)
num_seq_id num_seq_id_from_string(std::string const &s)
{
    static const std::unordered_map<std::string, num_seq_id> m{
        {"mersenne", num_seq_id::mersenne},       {"proth", num_seq_id::proth},
        {"riesel", num_seq_id::riesel},           {"smarandache", num_seq_id::smarandache},
        {"fibonacci", num_seq_id::fibonacci},     {"lucas", num_seq_id::lucas},
        {"proth_fixed_n", num_seq_id::proth_fixed_n}, {"riesel_fixed_n", num_seq_id::riesel_fixed_n},
        {"tribonacci", num_seq_id::tribonacci}};
    auto it = m.find(s);
    if (it == m.end()) throw std::runtime_error{"Invalid number sequence id `" + s + "'"};
    return it->second;
}

std::string to_string(num_seq_id id)
{
    static const char *names[] = {"mersenne", "proth", "riesel", "smarandache", "fibonacci",
                                  "lucas", "proth_fixed_n", "riesel_fixed_n", "tribonacci"};
    auto i = static_cast<uint32_t>(id);
    return i < 9 ? names[i] : "<unknown number sequence id>";
}

std::ostream &operator<<(std::ostream &out, trial_factoring_options const &o)
{
    return out << "n0 = " << o.n0 << ", n1 = " << o.n1 << ", f0 = " << o.f0 << ", f1 = " << o.f1 << ", output: `"
               << (o.output_path ? o.output_path->string() : std::string{"<none>"}) << "'";
}

// factors_buffer::report_excess_factors_if_any (factors_buffer.hpp:172-192)
template <>
void factors_buffer<uint64_t>::report_excess_factors_if_any(std::ostream &out)
{
    uint32_t prev = last_max_.value_or(0);
    uint32_t mx = counts_.empty() ? 0 : *std::max_element(counts_.begin(), counts_.end());
    last_max_ = mx;
    if (mx <= max_ || mx <= prev) return;
    if (!log().on(verbosity::normal)) return;
    uint32_t suggested = ((mx / 64) + 1) * 64 + 64;
    out << "At least one number has factors count " << mx << " which exceeds the current storage limit (" << max_
        << ").\n"
        << "Trial factoring continues normally, but some factors may be omitted from the recorded factor list for "
           "those terms.\n"
        << "Consider restarting with\n  --max-factors-per-number " << suggested << std::endl;
}

// ---------------------------------------------------------------- checkpoint v1

namespace {
constexpr uint32_t CHECKPOINT_VERSION = 1;

template <typename T>
void put(std::ostream &o, T const &x)
{
    o.write(reinterpret_cast<char const *>(&x), sizeof x);
}
template <typename T>
T get(std::istream &i)
{
    T x{};
    i.read(reinterpret_cast<char *>(&x), sizeof x);
    return x;
}
void put_str(std::ostream &o, std::string const &s)
{
    put(o, static_cast<std::string::size_type>(s.size()));
    o.write(s.data(), (std::streamsize)s.size());
}
std::string get_str(std::istream &i)
{
    auto n = get<std::string::size_type>(i);
    std::string s(n, '\0');
    i.read(s.data(), (std::streamsize)n);
    return s;
}
} // namespace

checkpoint_manager::checkpoint_manager(fs::path path, std::chrono::seconds period, job_spec spec)
    : path_(std::move(path)), period_(period), spec_(std::move(spec)), last_write_(std::chrono::steady_clock::now())
{
}

bool checkpoint_manager::due() const { return std::chrono::steady_clock::now() - last_write_ >= period_; }

// Layout (checkpoint.cpp:185-190, :91-115): u32 version, u32 seq_id, string
// seq_params, u64 n0 n1 f0 f1, string output_path, u32 max_factors, u64 last
// prime, then the factors buffer: u64 count, u32 max, u64 n0 and per index a
// u32 count followed by min(count, max) u64 factors.  Written to a
// temporary file and renamed over the checkpoint.
void checkpoint_manager::write(engine_state const &st)
{
    auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
                  std::chrono::system_clock::now().time_since_epoch())
                  .count();
    fs::path tmp = path_;
    tmp += ".tmp_";
    tmp += std::to_string(ns);
    {
        std::ofstream o{tmp, std::ios::binary};
        if (!o) throw std::runtime_error{"Could not create temporary file `" + tmp.string() + "'"};
        put(o, CHECKPOINT_VERSION);
        put(o, static_cast<uint32_t>(spec_.seq_spec.seq_id));
        put_str(o, spec_.seq_spec.seq_params);
        auto const &t = spec_.tf_options;
        put(o, t.n0);
        put(o, t.n1);
        put(o, t.f0);
        put(o, t.f1);
        put_str(o, t.output_path ? t.output_path->string() : std::string{});
        put(o, t.max_factors_per_number);
        put(o, st.last_processed_prime);
        auto const &fb = st.factors_buffer.get();
        put(o, fb.numbers_count());
        put(o, fb.max_factors_per_number());
        put(o, fb.n0());
        for (uint64_t i = 0; i < fb.numbers_count(); ++i) {
            uint32_t c = fb.counts()[i];
            put(o, c);
            uint64_t k = std::min(c, fb.max_factors_per_number());
            o.write(reinterpret_cast<char const *>(&fb.factors()[i * fb.max_factors_per_number()]),
                    (std::streamsize)(k * sizeof(uint64_t)));
        }
        o.flush();
    }
    fs::rename(tmp, path_);
    last_write_ = std::chrono::steady_clock::now();
}

std::optional<checkpoint_data> checkpoint_manager::load(fs::path const &path)
{
    std::ifstream in{path, std::ios::binary};
    if (!in) return {};
    auto version = get<uint32_t>(in);
    if (version != CHECKPOINT_VERSION)
        throw std::runtime_error{"Unexpected version number `" + std::to_string(version) +
                                 "' found in checkpoint file"};
    num_seq_spec spec{static_cast<num_seq_id>(get<uint32_t>(in)), get_str(in)};
    trial_factoring_options t;
    t.n0 = get<uint64_t>(in);
    t.n1 = get<uint64_t>(in);
    t.f0 = get<uint64_t>(in);
    t.f1 = get<uint64_t>(in);
    auto outp = get_str(in);
    if (!outp.empty()) t.output_path = fs::path{outp};
    t.max_factors_per_number = get<uint32_t>(in);
    auto last = get<uint64_t>(in);
    auto count = get<uint64_t>(in);
    auto max = get<uint32_t>(in);
    auto n0 = get<uint64_t>(in);
    factors_buffer<uint64_t> fb{n0, count, max};
    for (uint64_t i = 0; i < count; ++i) {
        uint32_t c = get<uint32_t>(in);
        fb.counts()[i] = c;
        uint64_t k = std::min(c, max);
        in.read(reinterpret_cast<char *>(&fb.factors()[i * max]), (std::streamsize)(k * sizeof(uint64_t)));
    }
    return checkpoint_data{job_spec{spec, t}, engine_state{last, factors_buffer_holder{std::move(fb)}}};
}

void checkpoint_manager::remove_checkpoint() { fs::remove(path_); }

// ---------------------------------------------------------------- device seam

namespace {

struct cb_state {
    trial_factoring_context *ctx;
    factors_buffer<uint64_t> *buf;
    progress *prog;
};

extern "C" void cb_progress(void *u, uint64_t bound, uint64_t last)
{
    auto *s = static_cast<cb_state *>(u);
    if (s->prog && last) s->prog->update(bound, last);
}
extern "C" int cb_due(void *u)
{
    auto *s = static_cast<cb_state *>(u);
    return s->ctx->checkpoint && s->ctx->checkpoint->due();
}
extern "C" void cb_checkpoint(void *u, uint64_t last)
{
    auto *s = static_cast<cb_state *>(u);
    s->buf->report_excess_factors_if_any(s->ctx->output_stream);
    s->ctx->checkpoint->write(engine_state{last, factors_buffer_holder{s->buf}});
}

} // namespace

void device_trial_factor_raw(trial_factoring_context &ctx, uint32_t seq_id, uint64_t param)
{
    auto const &opts = ctx.options;
    auto const &rt = ctx.runtime_options;
    auto &out = ctx.output_stream;
    ctd_hip_engine *eng = nullptr;
    if (ctd_hip_open(rt.device_id, &eng) != CTD_OK) {
        out << "Error selecting device " << rt.device_id << ": " << ctd_hip_last_error(nullptr) << std::endl;
        return;
    }
    std::unique_ptr<ctd_hip_engine, void (*)(ctd_hip_engine *)> guard{eng, ctd_hip_close};
    {
        char name[256] = {0};
        uint64_t mem = 0;
        int32_t cus = 0;
        ctd_hip_device_info(eng, name, sizeof name, &mem, &cus);
        out << "Using GPU device " << rt.device_id << ": " << name << " (" << std::setprecision(3)
            << double(mem) / (1024.0 * 1024 * 1024) << " GiB)\n";
    }
    std::unique_ptr<factors_buffer<uint64_t>> owned;
    factors_buffer<uint64_t> *buf;
    if (ctx.resume_state) {
        buf = &ctx.resume_state->factors_buffer.get();
        out << "Resuming at the smallest prime > " << ctx.resume_state->last_processed_prime << "." << std::endl;
    } else {
        owned = std::make_unique<factors_buffer<uint64_t>>(opts.n0, opts.n1 - opts.n0, opts.max_factors_per_number);
        buf = owned.get();
    }
    uint64_t f0 = ctx.resume_state ? ctx.resume_state->last_processed_prime + 1 : opts.f0;
    std::unique_ptr<progress> prog;
    if (rt.is_progress_enabled) prog = std::make_unique<progress>(f0, opts.f1, rt.progress_period, out);
    out << "Sieving for primes on the device." << std::endl;

    ctd_job job{seq_id, opts.max_factors_per_number, param, opts.n0, opts.n1, f0, opts.f1};
    ctd_run_opts ro{};
    // a device batch below 2^31 integers leaves most CUs idle during the sieve
    ro.batch_length = std::max<uint64_t>(rt.segment_length, uint64_t{1} << 31);
    ro.grid_size = rt.grid_size.value_or(0);
    ro.block_size = rt.block_size.value_or(0);
    cb_state st{&ctx, buf, prog.get()};
    ctd_callbacks cb{&st, cb_progress, cb_due, cb_checkpoint};
    uint64_t last = 0;
    int rc = ctd_hip_trial_factor(eng, &job, &ro, &cb, buf->counts().data(), buf->factors().data(), &last);
    if (prog) prog->end();
    if (rc != CTD_OK) {
        out << "Error executing trial division kernel" << std::endl
            << "Launch error: " << ctd_hip_last_error(eng) << std::endl;
    }
    ctd_run_stats stats{};
    ctd_hip_last_stats(eng, &stats);
    buf->report_excess_factors_if_any(out);
    ctx.results = buf->to_factoring_results<uint64_t, uint32_t>();
    out << "[Factoring took " << stats.total_ms / 1000 << "s (sieve: " << stats.sieve_ms / 1000 << "s)]"
        << std::endl;
}

// ---------------------------------------------------------------- engine front

void trial_factor_front_begin(trial_factoring_context &ctx, bool &proceed, std::unique_ptr<std::ostream> &file)
{
    auto const &o = ctx.options;
    auto &out = ctx.output_stream;
    proceed = false;
    if (o.n1 <= o.n0) {
        out << "Range [" << o.n0 << ", " << o.n1 << "[ is empty, no number to process..." << std::endl;
        return;
    }
    if (o.f1 <= o.f0) {
        out << "Range [" << o.f0 << ", " << o.f1 << "[ is empty, no prime numbers to divide by..." << std::endl;
        return;
    }
    if (o.output_path) {
        if (!ctx.checkpoint && fs::exists(*o.output_path))
            throw std::runtime_error{"Output: cannot overwrite existing file " + o.output_path->string()};
        file = std::make_unique<std::ofstream>(*o.output_path);
    }
    out << "Trial factoring starts, results will be written to "
        << (file ? std::string{"file `"} + o.output_path->string() + "'" : std::string{"the console."}) << std::endl;
    out << "TF range is [" << o.f0 << ", " << o.f1 << "[." << std::endl;
    if (ctx.checkpoint) {
        out << "Checkpoint file will be written to `" << ctx.checkpoint->checkpoint_path().string() << "' (period is "
            << ctx.checkpoint->period().count() << "s)." << std::endl;
        out << "Use `--resume " << ctx.checkpoint->checkpoint_path() << "' to resume execution." << std::endl;
    }
    proceed = true;
}

void trial_factor_front_end(trial_factoring_context &ctx, std::unique_ptr<std::ostream> &file)
{
    (file ? *file : ctx.output_stream) << ctx.results;
    if (file) file->flush();
    if (ctx.checkpoint) ctx.checkpoint->remove_checkpoint();
    if (log().on(verbosity::normal)) {
        ctx.output_stream << "Maximum number of factors found for a single sequence term: "
                          << ctx.results.max_factors_count() << std::endl
                          << "Overall number of factors found: " << ctx.results.factors_count() << std::endl;
    }
}

// ---------------------------------------------------------------- public API

namespace {
factoring_results<uint64_t, uint32_t> run(num_seq_spec spec, trial_factoring_options const &opts,
                                          tf_runtime_options const &rt, std::ostream &out,
                                          checkpoint_manager *ck, engine_state *resume)
{
    factoring_results<uint64_t, uint32_t> results{opts.n0, opts.n1 > opts.n0 ? opts.n1 - opts.n0 : 0};
    dispatch_num_seq<uint64_t>(spec, [&]<typename Seq, typename... Args>(Args &&...args) {
        trial_factor<Seq>(trial_factoring_context{opts, rt, results, out, ck, resume}, std::forward<Args>(args)...);
    });
    return results;
}
} // namespace

factoring_results<uint64_t, uint32_t> trial_factor(num_seq_spec spec, trial_factoring_options const &opts,
                                                  tf_runtime_options const &rt, std::ostream &out)
{
    std::unique_ptr<checkpoint_manager> ck;
    if (opts.output_path)
        ck = std::make_unique<checkpoint_manager>(fs::path{*opts.output_path} += ".chkpnt", rt.checkpoint_period,
                                                  job_spec{spec, opts});
    return run(spec, opts, rt, out, ck.get(), nullptr);
}

factoring_results<uint64_t, uint32_t> resume_trial_factoring(fs::path const &path, tf_runtime_options const &rt,
                                                             std::ostream &out)
{
    out << "Resuming from checkpoint `" << path.string() << "'..." << std::endl;
    auto data = checkpoint_manager::load(path);
    if (!data) throw std::runtime_error{"Failed to load checkpoint data from file `" + path.string() + "'"};
    auto ckpath = fs::path{data->spec.tf_options.output_path.value()} += ".chkpnt";
    checkpoint_manager ck{ckpath, rt.checkpoint_period, data->spec};
    return run(data->spec.seq_spec, data->spec.tf_options, rt, out, &ck, &data->state);
}

} // namespace cutrialdive
