// ctd_tf_fast.hpp -- narrow per-family index loops for primes below 2^63.
//
// Set-up (powers, doubling formulas, jump-ahead) is the Montgomery code of
// ctd_families.hpp; the per-index step then uses the cheapest exact form for
// the family.  A wave whose primes are all below 2^31 runs the doubling and
// addition families on 32-bit residues.
#pragma once

#include "ctd_families.hpp"

namespace ctd {

// One atomic per wave for the lanes whose residue is zero at index idx
// (device_factors_buffer.hpp:119-129 does one atomic per hit).  All active
// lanes of the wave are at the same index when this is called.
__device__ __forceinline__ void record_hits(const JobArgs &J, uint64_t idx, uint64_t p, bool hit)
{
    const uint64_t mask = __ballot(hit);
    if (mask) {
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&J.counts[idx], (uint32_t)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (hit) {
            uint32_t pos = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
            if (pos < J.maxf) J.factors[idx * J.maxf + pos] = p;
        }
    }
}

constexpr uint64_t FAST_PMAX = 1ull << 63;

// True when every lane of the wave holds a prime below 2^31.
__device__ __forceinline__ bool wave_w32(uint64_t p) { return __ballot(p >= (1ull << 31)) == 0; }

template <uint32_t FAM>
struct Fast {
    __host__ __device__ static bool fits(uint64_t pmax) { return pmax < FAST_PMAX; }

    __device__ static void run(const JobArgs &J, uint64_t p, uint64_t ns, uint64_t ne)
    {
        Mont M;
        M.init(p);
        const uint64_t steps = ne - ns - 1;
        uint64_t idx = ns - J.n0;
        if constexpr (FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL) {
            // x = c 2^n, hit when x == target; step x = 2x mod p
            uint64_t x = M.pow2(ns);
            x = (FAM == FAM_MERSENNE) ? M.to_plain(x) : M.mul(J.param % p, x);
            const uint64_t tgt = (FAM == FAM_PROTH) ? p - 1 : 1;
            record_hits(J, idx, p, x == tgt);
            if (wave_w32(p)) {
                const uint32_t q = (uint32_t)p, t = (uint32_t)tgt;
                uint32_t y = (uint32_t)x;
                for (uint64_t k = 0; k < steps; ++k) {
                    y <<= 1;
                    y = min(y, y - q);
                    ++idx;
                    record_hits(J, idx, p, y == t);
                }
            } else {
                for (uint64_t k = 0; k < steps; ++k) {
                    uint64_t u = x << 1;
                    uint64_t v = u - p;
                    x = (u >= p) ? v : u;
                    ++idx;
                    record_hits(J, idx, p, x == tgt);
                }
            }
        } else if constexpr (FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N) {
            // x = S(k); step x += 2^n mod p; hit x == 0
            const uint64_t cM = M.pow2(J.param);
            const uint64_t c = M.to_plain(cM);
            uint64_t x = M.mul(ns % p, cM);
            if constexpr (FAM == FAM_PROTH_FIXED_N)
                x = (x == p - 1) ? 0 : x + 1;
            else
                x = x ? x - 1 : p - 1;
            record_hits(J, idx, p, x == 0);
            if (wave_w32(p)) {
                const uint32_t q = (uint32_t)p, cc = (uint32_t)c;
                uint32_t y = (uint32_t)x;
                for (uint64_t k = 0; k < steps; ++k) {
                    y += cc;
                    y = min(y, y - q);
                    ++idx;
                    record_hits(J, idx, p, y == 0);
                }
            } else {
                const uint64_t cn = c - p; // x + c - p, modulo 2^64
                for (uint64_t k = 0; k < steps; ++k) {
                    uint64_t v = x + cn;           // wraps when x + c < p
                    x = (v < x) ? v : x + c;       // carry out <=> x + c >= p
                    ++idx;
                    record_hits(J, idx, p, x == 0);
                }
            }
        } else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS || FAM == FAM_TRIBONACCI ||
                             FAM == FAM_SMARANDACHE || FAM == FAM_PEANO) {
            GenState s;
            Gen<FAM>::init(s, M, J, ns);
            record_hits(J, idx, p, Gen<FAM>::zero(s, M));
            if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) {
                if (wave_w32(p)) {
                    const uint32_t q = (uint32_t)p;
                    uint32_t a = (uint32_t)s.x, b = (uint32_t)s.y;
                    for (uint64_t k = 0; k < steps; ++k) {
                        uint32_t t = a + b;
                        t = min(t, t - q);
                        a = b;
                        b = t;
                        ++idx;
                        record_hits(J, idx, p, b == 0);
                    }
                } else {
                    uint64_t a = s.x, b = s.y;
                    for (uint64_t k = 0; k < steps; ++k) {
                        uint64_t t = a + b;
                        uint64_t v = t - p;
                        t = (t >= p) ? v : t;
                        a = b;
                        b = t;
                        ++idx;
                        record_hits(J, idx, p, b == 0);
                    }
                }
            } else if constexpr (FAM == FAM_TRIBONACCI) {
                if (wave_w32(p)) {
                    const uint32_t q = (uint32_t)p;
                    uint32_t a = (uint32_t)s.x, b = (uint32_t)s.y, c = (uint32_t)s.z;
                    for (uint64_t k = 0; k < steps; ++k) {
                        uint32_t t = a + b;
                        t = min(t, t - q);
                        t += c;
                        t = min(t, t - q);
                        a = b;
                        b = c;
                        c = t;
                        ++idx;
                        record_hits(J, idx, p, a == 0);
                    }
                } else {
                    uint64_t a = s.x, b = s.y, c = s.z;
                    for (uint64_t k = 0; k < steps; ++k) {
                        uint64_t t = a + b;
                        t = (t >= p) ? t - p : t;
                        t += c;
                        t = (t >= p) ? t - p : t;
                        a = b;
                        b = c;
                        c = t;
                        ++idx;
                        record_hits(J, idx, p, a == 0);
                    }
                }
            } else {
                for (uint64_t n = ns; n + 1 < ne; ++n) {
                    Gen<FAM>::step(s, M, n);
                    ++idx;
                    record_hits(J, idx, p, Gen<FAM>::zero(s, M));
                }
            }
        }
    }
};

} // namespace ctd
