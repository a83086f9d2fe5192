// ctd_ranges.hpp -- overflow-safe arithmetic on prime windows.
//
// Prime windows run up to f1 = 2^64 - 1 (--tf-bits 64 means [2, 2^64 - 1),
// command_line.cpp:317-325), so the end of a segment, batch or tile that
// starts near the top of the range cannot be written as lo + len: the sum
// wraps.  Every such bound goes through these functions instead.
//
// Plain constexpr C++ with no HIP dependency: the device sieve, the host
// schedulers and a g++ unit test include the same code (clang compiles
// constexpr functions for the device as well).
#pragma once

#include <stdint.h>

namespace ctd {

constexpr uint64_t U64_MAX = ~0ull;

// a + b, or 2^64 - 1 when the sum does not fit.
constexpr uint64_t sat_add(uint64_t a, uint64_t b) { return a > U64_MAX - b ? U64_MAX : a + b; }

// End of the block that starts at lo < end: min(lo + len, end), without
// forming lo + len when it would pass end.
constexpr uint64_t block_end(uint64_t lo, uint64_t len, uint64_t end) { return end - lo > len ? lo + len : end; }

// Largest integer a sieve segment [lo, lo + span) has to decide when only
// numbers below end are emitted: min(lo + span, end) - 1, for lo < end.
// Sieving numbers d with d*d above it mark nothing in the segment.
constexpr uint64_t seg_last(uint64_t lo, uint64_t span, uint64_t end) { return block_end(lo, span, end) - 1; }

// floor(sqrt(n)) for every 64-bit n, in integers only (bit by bit from the
// top; the candidate stays below 2^32, so its square never wraps).
constexpr uint64_t isqrt_u64(uint64_t n)
{
    uint64_t r = 0;
    for (int b = 31; b >= 0; --b) {
        const uint64_t c = r | (1ull << b);
        if (c * c <= n) r = c;
    }
    return r;
}

// len rounded up to a whole number of spans, at least one span and at most
// max_spans of them (a requested length near 2^64 would otherwise round up
// to 0 and the block loop would never advance).
constexpr uint64_t whole_spans(uint64_t len, uint64_t span, uint64_t max_spans)
{
    uint64_t k = len / span + (len % span != 0);
    if (k < 1) k = 1;
    if (k > max_spans) k = max_spans;
    return k * span;
}

} // namespace ctd
