// ranges_tool.cpp -- evaluates the window arithmetic of ctd_ranges.hpp on
// the host, for tests/test_host.py, which checks every answer with Python's
// exact integers.  Each input line is "<op> <a> <b> <c>"; each output line
// is the result.  Ops: isqrt a | sat_add a b | block_end a b c |
// seg_last a b c | whole_spans a b c | blocks lo0 len end (the number of
// blocks a loop of block_end takes from lo0 to end, and their total length).
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "ctd_ranges.hpp"

int main()
{
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string op;
        uint64_t a = 0, b = 0, c = 0;
        in >> op >> a >> b >> c;
        if (op == "isqrt") std::printf("%llu\n", (unsigned long long)ctd::isqrt_u64(a));
        else if (op == "sat_add") std::printf("%llu\n", (unsigned long long)ctd::sat_add(a, b));
        else if (op == "block_end") std::printf("%llu\n", (unsigned long long)ctd::block_end(a, b, c));
        else if (op == "seg_last") std::printf("%llu\n", (unsigned long long)ctd::seg_last(a, b, c));
        else if (op == "whole_spans") std::printf("%llu\n", (unsigned long long)ctd::whole_spans(a, b, c));
        else if (op == "blocks") {
            uint64_t n = 0, covered = 0;
            for (uint64_t lo = a; lo < c; ++n) { // the loop shape of the segment, batch and tile loops
                const uint64_t hi = ctd::block_end(lo, b, c);
                if (hi <= lo) return 2; // would never advance
                covered += hi - lo;
                lo = hi;
            }
            std::printf("%llu %llu\n", (unsigned long long)n, (unsigned long long)covered);
        } else return 1;
    }
    return 0;
}
