// ctd_cli.cpp -- the `cutrialdive` command line (command_line.cpp:223-373,
// main.cpp:95-146) over the HIP engine.  Same options, same order rules,
// same messages and exit codes; the trial factoring runs on the GPU.
#include <array>
#include <cstring>
#include <iostream>
#include <sstream>

#include "ctd_api.hpp"

namespace ctd = cutrialdive;
using namespace cutrialdive;

namespace {

constexpr char const *VERSION = "0.4.0-mi355x";

struct parse_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct cli_options {
    bool is_resuming = false;
    std::optional<std::filesystem::path> checkpoint_path;
    std::optional<uint32_t> checkpoint_period;
    std::optional<bool> is_progress_enabled;
    std::optional<uint32_t> threads_count;
    std::optional<uint64_t> segment_length;
    std::optional<verbosity> verbosity_level;
    std::optional<int32_t> device_id, grid_size, block_size;
    num_seq_id seq_id{};
    std::string seq_params;
    bool wants_value = false, wants_expression = false;
    uint64_t n = 0;
    std::optional<trial_factoring_options> tf_options;
};

// `flag value` at *argv: reads the value with operator>> (so "12abc" reads 12
// and "abc" reads 0, as std::istringstream does in command_line.cpp:23-35)
// and leaves argv on the value.
template <typename T, typename Act>
bool eat(char const *flag, char **&argv, Act act)
{
    if (!*argv || std::strcmp(*argv, flag)) return false;
    ++argv;
    if (!*argv) throw parse_error{std::string{"Missing value for option "} + flag};
    if constexpr (std::is_same_v<T, std::string>) {
        act(std::string{*argv});
    } else {
        T v{};
        std::istringstream{*argv} >> v;
        act(v);
    }
    return true;
}
template <typename T, typename D>
bool eat_into(char const *flag, D &dest, char **&argv)
{
    return eat<T>(flag, argv, [&](T v) { dest = v; });
}

bool parse_no_progress(cli_options &o, char **argv)
{
    if (std::strcmp(*argv, "--no-progress")) return false;
    o.is_progress_enabled = false;
    return true;
}

bool parse_verbosity(cli_options &o, char **&argv)
{
    return eat<uint16_t>("--verbosity", argv, [&](uint16_t v) {
        if (!is_valid_verbosity(v)) throw parse_error{"Invalid value for option --verbosity"};
        o.verbosity_level = static_cast<verbosity>(v);
    });
}

bool parse_segment_length(cli_options &o, char **&argv)
{
    if (eat<uint64_t>("--segment-length", argv, [&](uint64_t len) {
            if (o.segment_length)
                throw parse_error{"Invalid command line, --segment-length is not allowed when --segment-bits is present"};
            o.segment_length = len;
        }))
        return true;
    return eat<int32_t>("--segment-bits", argv, [&](int32_t bits) {
        if (o.segment_length)
            throw parse_error{"Invalid command line, --segment-bits is not allowed when --segment-length is present"};
        if (bits < 0 || bits >= 64) throw parse_error{"value of option --segment-bits must be an integer in range [0, 63]"};
        o.segment_length = uint64_t{1} << bits;
    });
}

bool parse_device_options(cli_options &o, char **&argv)
{
    return eat_into<int32_t>("--device", o.device_id, argv) || eat_into<int32_t>("--grid-size", o.grid_size, argv) ||
           eat_into<int32_t>("--block-size", o.block_size, argv);
}

void parse_num_seq(cli_options &o, char **&argv)
{
    bool ok = eat<std::string>("--num-seq", argv, [&](std::string spec) {
        try {
            auto comma = spec.find(',');
            if (comma != std::string::npos) {
                o.seq_id = num_seq_id_from_string(spec.substr(0, comma));
                o.seq_params = spec.substr(comma + 1);
                if (o.seq_params.empty()) throw parse_error{"Invalid value for option --num-seq"};
            } else {
                o.seq_id = num_seq_id_from_string(spec);
            }
        } catch (std::exception const &e) {
            throw parse_error{e.what()};
        }
        ++argv;
    });
    if (!ok) throw parse_error{"Invalid command line, option --num-seq must come first"};
}

// Resume mode (command_line.cpp:166-208): the options allowed next to
// --resume, in any order; returns at the first other option when not resuming.
void parse_resume(cli_options &o, char **&argv)
{
    bool have_period = false;
    for (; *argv; ++argv) {
        if (eat<std::string>("--resume", argv, [&](std::string p) {
                o.is_resuming = true;
                o.checkpoint_path = p;
            }))
            continue;
        if (parse_device_options(o, argv)) continue;
        if (eat<uint32_t>("--checkpoint-period", argv, [&](uint32_t s) {
                o.checkpoint_period = s;
                have_period = true;
            }))
            continue;
        if (parse_verbosity(o, argv)) continue;
        if (eat_into<uint32_t>("--threads", o.threads_count, argv)) continue;
        if (parse_no_progress(o, argv)) continue;
        if (parse_segment_length(o, argv)) continue;
        if (!o.is_resuming && !have_period) return;
    }
    if (have_period && !o.is_resuming) throw parse_error{"Invalid command line, option --checkpoint-period is unexpected"};
}

enum class request { usage, help, version, autotest, run };

request parse(int argc, char **argv, cli_options &o)
{
    if (argc == 1) return request::usage;
    if (argc == 2) {
        if (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) return request::help;
        if (!std::strcmp(argv[1], "--version")) return request::version;
        if (!std::strcmp(argv[1], "--autotest")) return request::autotest;
    }
    char **a = argv + 1;
    parse_resume(o, a);
    if (o.is_resuming) return request::run;
    if (parse_verbosity(o, a)) ++a;
    parse_num_seq(o, a);

    trial_factoring_options tf{};
    bool have_start = false, have_end = false, have_bits = false, have_tf_start = false, have_tf_end = false;
    for (; *a; ++a) {
        if (parse_verbosity(o, a)) continue;
        if (eat_into<uint32_t>("--threads", o.threads_count, a)) continue;
        if (eat<uint64_t>("--print", a, [&](uint64_t n) {
                o.wants_value = true;
                o.n = n;
            }))
            return request::run;
        if (eat<uint64_t>("--print-expression", a, [&](uint64_t n) {
                o.wants_expression = true;
                o.n = n;
            }))
            return request::run;
        if (parse_device_options(o, a)) continue;
        if (eat_into<uint64_t>("-s", tf.n0, a)) {
            have_start = true;
            continue;
        }
        if (eat_into<uint64_t>("-e", tf.n1, a)) {
            have_end = true;
            continue;
        }
        if (eat<uint16_t>("--tf-bits", a, [&](uint16_t bits) {
                tf.f0 = 2;
                if (bits > 64) throw parse_error{"The value of option --tf-bits must be an integer between 0 and 64"};
                tf.f1 = bits < 64 ? uint64_t{1} << bits : ~uint64_t{};
            })) {
            have_bits = true;
            continue;
        }
        if (eat_into<uint64_t>("--tf-start", tf.f0, a)) {
            have_tf_start = true;
            continue;
        }
        if (eat_into<uint64_t>("--tf-end", tf.f1, a)) {
            have_tf_end = true;
            continue;
        }
        if (have_bits && (have_tf_start || have_tf_end))
            throw parse_error{"Mixing --tf-bits with --tf-start and --tf-end is not allowed"};
        if (parse_no_progress(o, a)) continue;
        if (eat<std::string>("--output", a, [&](std::string p) { tf.output_path = std::filesystem::path{p}; }))
            continue;
        if (eat_into<uint32_t>("--checkpoint-period", o.checkpoint_period, a)) continue;
        if (parse_segment_length(o, a)) continue;
        if (eat_into<uint32_t>("--max-factors-per-number", tf.max_factors_per_number, a)) continue;
        // anything else is skipped, as the reference parser does
    }
    if (!have_start || !(have_tf_start || have_bits) || !(have_tf_end || have_bits))
        throw parse_error{"Missing mandatory option"};
    if (!have_end) tf.n1 = tf.n0 + 1;
    if (tf.output_path && tf.output_path->empty()) throw parse_error{"Invalid output path"};
    o.tf_options = std::move(tf);
    return request::run;
}

void print_version(std::ostream &out)
{
    out << "cutrialdive version " << VERSION << std::endl
        << "[Compiled with gcc " << __VERSION__ << "] [engine: HIP gfx950, C ABI " << ctd_hip_abi_version() << "]"
        << std::endl;
}

void print_usage_only(std::ostream &out)
{
    out << R"(
  cutrialdive ( --version
              | ( -h | --help )
              | ( --num-seq <NUM_SEQ_ID>
                        ( --print n
                        | --print-expression n
                        | <trial factoring options>
                        )
                )
              | --resume <checkpoint_path> <resume options>
              )
              [<global options>]
)";
}

void print_usage(std::ostream &out)
{
    print_usage_only(out);
    out << "\n\n  Type 'cutrialdive --help' for more.\n";
}

void print_help(std::ostream &out)
{
    print_usage_only(out);
    out << R"(
    --help                 prints this help and exits.
    --version              prints the version number and exits.
    --autotest             runs the device self tests (sieve windows with known primes).

    Global options
    --threads <N>          accepted for compatibility; the engine runs on the GPU.
    --verbosity <V>        log level in [0, 4]: silent, normal, verbose, debug, diagnostic.

    Number sequence selection
    --num-seq <NUM_SEQ_ID> mersenne, proth,k, riesel,k, smarandache[,base] (base 2..64,
                           default 10), fibonacci, lucas, proth_fixed_n,n,
                           riesel_fixed_n,n, tribonacci.

    Print number
    --print n              prints S(n) in base 10.
    --print-expression n   prints an expression whose evaluation yields S(n).

    Trial factoring (indices [n0, n1[, primes [f0, f1[)
    -s n0                  first index.
    -e n1                  end index (excluded), default n0 + 1.
    --tf-bits b            primes in [2, 2^b[.
    --tf-start f0          lower bound of the primes.
    --tf-end f1            upper bound of the primes (excluded).
    --no-progress          suppresses progress display.
    --output <path>        writes the results to path; path.chkpnt is the checkpoint
                           file, deleted when the run completes.
    --checkpoint-period T  minimum seconds between two checkpoint savings.
    --segment-bits <L>     device batch of 2^L integers (at least 2^31 are used).
    --segment-length <N>   device batch of N integers (at least 2^31 are used).
    --max-factors-per-number <N>  factors stored per term (default 32); more are
                           counted and reported.
    --device <id>          GPU to run on.
    --grid-size N          workgroups per launch of the residue kernel.
    --block-size N         threads per workgroup of the residue kernel.

    Resuming
    --resume <checkpoint file path>  resumes an interrupted trial factoring; allowed
                           next to it: --threads, --no-progress, --checkpoint-period,
                           --device, --grid-size, --block-size, --segment-length,
                           --segment-bits, --verbosity.

    Examples
    cutrialdive --num-seq proth,12345678910111213 -s 30000000 -e 30050000 --tf-bits 40 --output output.txt
    cutrialdive --resume output.txt.chkpnt
    cutrialdive --num-seq smarandache,2 --print-expression 37
)";
}

// Known prime windows checked on the device (autotests.cpp:148-353 holds the
// same kind of checks for the host sieve).
int autotest()
{
    std::cout << "Executing self tests..." << std::endl;
    ctd_hip_engine *eng = nullptr;
    if (ctd_hip_open(0, &eng) != CTD_OK) {
        std::cerr << "Failure: " << ctd_hip_last_error(nullptr) << std::endl;
        return 1;
    }
    struct window {
        uint64_t lo, hi, count, first, last;
    };
    const std::array<window, 6> windows{{
        {256, 307, 8, 257, 293},
        {65536, 65537, 0, 0, 0},
        {65536, 65538, 1, 65537, 65537},
        {1ull << 32, (1ull << 32) + 1000, 56, 4294967311ull, 4294968289ull},
        {(1ull << 61) - 100, 1ull << 61, 3, 2305843009213693907ull, 2305843009213693951ull},
        {0, 1ull << 17, 12251, 2, 131071},
    }};
    int failures = 0;
    std::vector<uint64_t> buf(1 << 15);
    for (auto const &w : windows) {
        uint64_t n = 0;
        int rc = ctd_hip_sieve(eng, w.lo, w.hi, buf.data(), buf.size(), &n);
        bool ok = rc == CTD_OK && n == w.count && (!n || (buf[0] == w.first && buf[n - 1] == w.last));
        if (!ok) {
            ++failures;
            std::cerr << "Failure: expected " << w.count << " primes in [" << w.lo << ", " << w.hi << "), got " << n
                      << std::endl;
        }
    }
    ctd_hip_close(eng);
    std::cout << (failures ? "Self tests failed." : "All self tests passed.") << std::endl;
    return failures ? 1 : 0;
}

tf_runtime_options runtime_options_of(cli_options const &o)
{
    auto rt = tf_runtime_options::default_options();
    if (o.threads_count) rt.common_options.threads_count = *o.threads_count;
    if (o.verbosity_level) rt.common_options.verbosity_level = *o.verbosity_level;
    if (o.checkpoint_period) rt.checkpoint_period = std::chrono::seconds{*o.checkpoint_period};
    if (o.is_progress_enabled) rt.is_progress_enabled = *o.is_progress_enabled;
    if (o.segment_length) rt.segment_length = *o.segment_length;
    if (o.device_id) rt.device_id = *o.device_id;
    if (o.grid_size) rt.grid_size = *o.grid_size;
    if (o.block_size) rt.block_size = *o.block_size;
    return rt;
}

} // namespace

int main(int argc, char **argv)
{
    try {
        cli_options o;
        switch (parse(argc, argv, o)) {
        case request::version: print_version(std::cout); return 0;
        case request::usage: print_usage(std::cerr); return 1;
        case request::help: print_help(std::cout); return 0;
        case request::autotest: return autotest();
        case request::run: break;
        }
        auto rt = runtime_options_of(o);
        log().level = rt.common_options.verbosity_level;
        log().out = &std::cout;
        if (o.is_resuming) {
            ctd::resume_trial_factoring(o.checkpoint_path.value(), rt, std::cout);
            return 0;
        }
        num_seq_spec spec{o.seq_id, o.seq_params};
        if (o.wants_value || o.wants_expression) {
            return dispatch_num_seq<uint64_t>(spec, [&]<typename Seq, typename... Args>(Args &&...args) {
                auto m = get_math_sequence(Seq{std::forward<Args>(args)...});
                if (o.wants_value) {
                    if constexpr (requires { m.print_value(o.n, std::cout); })
                        m.print_value(o.n, std::cout) << std::endl;
                    else
                        std::cout << "Value not available" << std::endl;
                } else {
                    if constexpr (requires { m.print_expression(o.n, std::cout); })
                        m.print_expression(o.n, std::cout) << std::endl;
                    else
                        std::cout << "Expression not available" << std::endl;
                }
                return 0;
            });
        }
        ctd::trial_factor(spec, *o.tf_options, rt, std::cout);
    } catch (std::exception const &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
