// ctd_progress.hpp -- the progress line "[xx.xx% done][last prime][Elapsed ..][ETA ..]"
// (progress.cpp:16-113), updated at most once per period.
#pragma once

#include <chrono>
#include <cstdint>
#include <iomanip>
#include <ostream>
#include <string>

namespace cutrialdive {

class progress {
public:
    progress(uint64_t start, uint64_t target, std::chrono::milliseconds period, std::ostream &out)
        : start_(start), target_(target), period_(period), out_(out), t0_(std::chrono::steady_clock::now()),
          last_(t0_)
    {
        for (uint64_t t = target; t; t /= 10) ++digits_;
    }
    ~progress()
    {
        if (!ended_) out_ << std::endl;
    }
    void update(uint64_t current, uint64_t last_prime)
    {
        auto now = std::chrono::steady_clock::now();
        if (now - last_ < period_ && current != target_) return;
        double pct = target_ > start_ ? double(((unsigned __int128)(current - start_) * 10000000000ull) /
                                               (target_ - start_)) / 100000000.0
                                      : 100.0;
        double el = std::chrono::duration<double>(now - t0_).count();
        out_ << "\r" << std::string(128, ' ') << "\r[" << std::setw(6) << std::fixed << std::setprecision(2) << pct
             << "% done][" << std::setw(digits_) << last_prime << "][Elapsed ";
        duration(el);
        out_ << "]";
        if (current != target_ && pct > 0) {
            out_ << "[ETA ";
            duration((100 - pct) / pct * el);
            out_ << "]";
        }
        out_ << std::flush;
        last_ = std::chrono::steady_clock::now();
    }
    void end()
    {
        if (ended_) return;
        out_ << std::endl;
        ended_ = true;
    }

private:
    void duration(double s)
    {
        if (s > 365 * 86400.0) {
            out_ << std::setw(8) << std::fixed << std::setprecision(2) << s / (365 * 86400.0) << " years";
            return;
        }
        if (s > 86400) {
            out_ << std::setw(3) << int32_t(s / 86400) << "d ";
            s -= 86400 * int32_t(s / 86400);
        }
        int32_t h = int32_t(s / 3600), m = int32_t((s - 3600 * h) / 60), sec = int32_t(s - 3600 * h - 60 * m);
        out_ << std::setfill('0') << std::setw(2) << h << ":" << std::setw(2) << m << ":" << std::setw(2) << sec
             << std::setfill(' ');
    }
    uint64_t start_, target_;
    std::chrono::milliseconds period_;
    std::ostream &out_;
    std::chrono::steady_clock::time_point t0_, last_;
    int digits_ = 0;
    bool ended_ = false;
};

} // namespace cutrialdive
