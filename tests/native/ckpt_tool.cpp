// ckpt_tool.cpp -- drives checkpoint_manager (cutrialdive_amd/host/ctd_host.cpp)
// from the checkpoint tests.
//   ckpt_tool write <path> <seq_id> <params> <n0> <n1> <f0> <f1> <out|-> <max> <last>
//       writes a checkpoint whose factors buffer holds count(i) = (5 i + 3) mod (max + 3)
//       and factor j of index i = 1000003 i + 17 j (only min(count, max) are stored)
//   ckpt_tool dump <path>
//       loads a checkpoint and prints every field, one per line
#include <cstdlib>
#include <iostream>

#include "ctd_api.hpp"

using namespace cutrialdive;

int main(int argc, char **argv)
{
    try {
        std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "write" && argc == 12) {
            trial_factoring_options t;
            t.n0 = std::strtoull(argv[5], nullptr, 10);
            t.n1 = std::strtoull(argv[6], nullptr, 10);
            t.f0 = std::strtoull(argv[7], nullptr, 10);
            t.f1 = std::strtoull(argv[8], nullptr, 10);
            if (std::string{argv[9]} != "-") t.output_path = std::filesystem::path{argv[9]};
            t.max_factors_per_number = (uint32_t)std::strtoul(argv[10], nullptr, 10);
            num_seq_spec spec{static_cast<num_seq_id>(std::strtoul(argv[3], nullptr, 10)), argv[4]};
            factors_buffer<uint64_t> fb{t.n0, t.n1 - t.n0, t.max_factors_per_number};
            for (uint64_t i = 0; i < t.n1 - t.n0; ++i) {
                uint32_t c = (uint32_t)((5 * i + 3) % (t.max_factors_per_number + 3));
                fb.counts()[i] = c;
                for (uint32_t j = 0; j < std::min(c, t.max_factors_per_number); ++j)
                    fb.factors()[i * t.max_factors_per_number + j] = 1000003 * i + 17 * j;
            }
            checkpoint_manager m{argv[2], std::chrono::seconds{0}, job_spec{spec, t}};
            m.write(engine_state{std::strtoull(argv[11], nullptr, 10), factors_buffer_holder{&fb}});
            return 0;
        }
        if (mode == "dump" && argc == 3) {
            auto d = checkpoint_manager::load(argv[2]);
            if (!d) {
                std::cout << "missing" << std::endl;
                return 3;
            }
            auto const &t = d->spec.tf_options;
            std::cout << "seq_id " << static_cast<uint32_t>(d->spec.seq_spec.seq_id) << "\n"
                      << "params " << d->spec.seq_spec.seq_params << "\n"
                      << "n0 " << t.n0 << "\nn1 " << t.n1 << "\nf0 " << t.f0 << "\nf1 " << t.f1 << "\n"
                      << "output " << (t.output_path ? t.output_path->string() : std::string{"-"}) << "\n"
                      << "max " << t.max_factors_per_number << "\n"
                      << "last " << d->state.last_processed_prime << "\n";
            auto const &fb = d->state.factors_buffer.get();
            std::cout << "buffer " << fb.numbers_count() << " " << fb.max_factors_per_number() << " " << fb.n0()
                      << "\n";
            for (uint64_t i = 0; i < fb.numbers_count(); ++i) {
                uint32_t c = fb.counts()[i];
                std::cout << c << ":";
                for (uint32_t j = 0; j < std::min(c, fb.max_factors_per_number()); ++j)
                    std::cout << " " << fb.factors()[i * fb.max_factors_per_number() + j];
                std::cout << "\n";
            }
            return 0;
        }
        std::cerr << "usage: ckpt_tool write|dump ..." << std::endl;
        return 2;
    } catch (std::exception const &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
