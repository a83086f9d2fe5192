// ctd_sieve.hpp -- segmented sieve of Eratosthenes in LDS (gfx950).
//
// Replaces the host prime supply: siever::sieve over primesieve
// (siever.hpp:179-237) and the per-segment H2D copy of
// device_prime_data::copy_from_host (prime_data.hpp:165-196).
//
// One workgroup sieves one segment of SEG_SPAN integers [lo, lo+SEG_SPAN),
// lo even.  Bit i of the LDS bitmap stands for the odd number lo + 2i + 1.
// Each thread owns 16 consecutive bitmap words (512 odd numbers).
//
//   1. Wheel: the odd multiples of 3..31 are not marked one by one; each
//      thread builds its 16 words in registers from per-prime bit patterns
//      (a 64-bit pattern of p's multiples, shifted by the word's phase), so
//      the densest primes cost no LDS atomics at all.
//   2. Sieving numbers 37..SMALL_LIMIT: the phase of each (its first odd
//      multiple >= max(p^2, lo+1)) is computed once, by one thread, into LDS;
//      then one wave per number marks its multiples (64 lanes, stride 64 p).
//   3. Larger sieving numbers: one per lane, computing its own phase.
//   4. Compaction: per-thread popcounts and a workgroup prefix sum give each
//      prime its position; the primes are staged in LDS (the bitmap is
//      reused once every thread holds its words in registers) and written to
//      HBM by consecutive lanes, so every 64-byte line of the prime stream is
//      written whole.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ctd_ranges.hpp"

namespace ctd {

constexpr uint32_t SEG_BITS = 1u << 19;            // odd numbers per segment
constexpr uint64_t SEG_SPAN = 2ull * SEG_BITS;      // integers per segment
constexpr uint32_t SEG_WORDS = SEG_BITS / 32;       // 16384 words = 64 KiB of LDS
constexpr uint32_t SIEVE_THREADS = 1024;
constexpr uint32_t SIEVE_WAVES = SIEVE_THREADS / 64;
constexpr uint32_t WORDS_PER_THREAD = SEG_WORDS / SIEVE_THREADS;
constexpr uint32_t SMALL_LIMIT = 1u << 13;          // wave-cooperative marking below this
constexpr uint32_t WHEEL_MAX = 31;                  // odd sieving numbers <= this come from patterns
constexpr uint32_t SMALL_MAX_COUNT = 2048;          // sieving numbers in (WHEEL_MAX, SMALL_LIMIT) at most
constexpr uint32_t STAGE_ENTRIES = SEG_WORDS;       // u32 staging slots in the reused bitmap

enum SieveMode : uint32_t {
    SIEVE_COUNT = 0,     // per-segment prime counts only
    SIEVE_ORDERED32 = 1, // u32 primes at a given per-segment offset (sieving base)
    SIEVE_ORDERED64 = 2, // u64 primes at a given per-segment offset (ordered lists)
    SIEVE_DENSE = 3      // u64 primes at an atomically reserved offset (prime stream)
};

struct SieveArgs {
    uint64_t lo0;            // even base of segment 0
    uint64_t start, end;     // emit primes of [start, end) only
    const uint32_t *sp;      // sieving numbers (ascending; odd; d*d covers end-1)
    const uint64_t *sp_mu;   // floor(2^64 / sp[i]) for the remainder of lo
    uint32_t nsp;            // count of sieving numbers
    uint32_t nsmall;         // sp[0..nsmall) are < SMALL_LIMIT
    uint32_t mode;
    uint32_t nwheel;         // sp[0..nwheel) are <= WHEEL_MAX (marked by the wheel patterns)
    uint32_t *seg_counts;    // [segments] (COUNT, and written by all modes)
    const uint64_t *seg_off; // ORDERED*: output offset per segment
    uint32_t *out32;         // ORDERED32 output
    uint64_t *out64;         // ORDERED64 / DENSE output
    unsigned long long *total;   // DENSE: running count (atomic)
    unsigned long long *maxp;    // DENSE: largest prime emitted (atomic max)
    uint64_t cap;            // DENSE: capacity of out64
};

// lo mod p for p < 2^32 using mu = floor(2^64 / p) (q underestimates by <= 1).
__device__ __forceinline__ uint32_t mod_by_mu(uint64_t lo, uint32_t p, uint64_t mu)
{
    uint64_t q = __umul64hi(lo, mu);
    uint64_t r = lo - q * p;
    return (uint32_t)(r >= p ? r - p : r);
}

// Index of the first bit >= 0 to mark for sieving number p in segment lo:
// the smallest odd multiple of p that is >= max(p^2, lo+1).
__device__ __forceinline__ uint64_t first_index(uint64_t lo, uint32_t p, uint64_t mu)
{
    uint64_t pp = (uint64_t)p * p;
    if (pp >= lo + 1) return (pp - lo - 1) >> 1;
    uint32_t r = mod_by_mu(lo, p, mu);
    uint32_t t = p - 1 - r; // -(lo + 1) mod p, as r < p
    return (t & 1) ? ((uint64_t)t + p) >> 1 : t >> 1; // 2i = t (mod p)
}

// Phase of the wheel prime P in segment lo: the bit index (mod P) of its odd
// multiples, i.e. i with lo + 2i + 1 = 0 (mod P).
template <uint32_t P>
__device__ __forceinline__ uint32_t wheel_phase(uint64_t lo)
{
    const uint32_t r = (uint32_t)(lo % P);          // constant divisor: multiply-high sequence
    const uint32_t t = (2 * P - 1 - r) % P;         // -(lo + 1) mod P
    return (t & 1) ? (t + P) >> 1 : t >> 1;         // 2i = t (mod P)
}

// ORs the multiples of P into the thread's words w0..w0+WORDS_PER_THREAD-1.
// pat holds bits 0, P, 2P, ... below 64; the word w gets pat << o with o the
// phase relative to bit 32w, so its low 32 bits are the multiples in w.
template <uint32_t P>
__device__ __forceinline__ void wheel_or(uint32_t (&words)[WORDS_PER_THREAD], uint32_t phase, uint32_t w0)
{
    uint64_t pat = 0;
    for (uint32_t b = 0; b < 64; b += P) pat |= 1ull << b;
    // o = (phase - 32 w0) mod P; 32 w0 < 2^19 < 2^19 P
    uint32_t o = (phase + (P << 19) - 32 * w0) % P;
    constexpr uint32_t r32 = 32 % P;
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) {
        words[j] |= (uint32_t)(pat << o);
        const uint32_t t = o - r32;
        o = min(t, t + P); // (o - 32) mod P without a branch (unsigned wrap)
    }
}

__device__ __forceinline__ void wheel_words(uint32_t (&words)[WORDS_PER_THREAD], uint64_t lo, uint32_t w0)
{
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) words[j] = 0;
    wheel_or<3>(words, wheel_phase<3>(lo), w0);
    wheel_or<5>(words, wheel_phase<5>(lo), w0);
    wheel_or<7>(words, wheel_phase<7>(lo), w0);
    wheel_or<11>(words, wheel_phase<11>(lo), w0);
    wheel_or<13>(words, wheel_phase<13>(lo), w0);
    wheel_or<17>(words, wheel_phase<17>(lo), w0);
    wheel_or<19>(words, wheel_phase<19>(lo), w0);
    wheel_or<23>(words, wheel_phase<23>(lo), w0);
    wheel_or<29>(words, wheel_phase<29>(lo), w0);
    wheel_or<31>(words, wheel_phase<31>(lo), w0);
}

// Marks the bitmap of segment lo.  On return every thread holds its words
// (the complement gives the survivors) and the bitmap is no longer needed.
__device__ __forceinline__ void sieve_mark(uint32_t *bits, uint32_t *phase, const SieveArgs &A, uint64_t lo,
                                           uint32_t (&words)[WORDS_PER_THREAD])
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t w0 = tid * WORDS_PER_THREAD;
    // 1. wheel patterns (every odd multiple of 3..31, the primes themselves
    //    included; they are unmarked again below when the segment holds them)
    wheel_words(words, lo, w0);
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) bits[w0 + j] = words[j];
    if (lo < WHEEL_MAX && tid == 0) { // the words of thread 0 hold 3..31 here
        constexpr uint32_t wheel_primes = (1u << 3) | (1u << 5) | (1u << 7) | (1u << 11) | (1u << 13) |
                                          (1u << 17) | (1u << 19) | (1u << 23) | (1u << 29) | (1u << 31);
        for (uint32_t p = 3; p <= WHEEL_MAX; p += 2)
            if (p > lo && ((wheel_primes >> p) & 1)) {
                const uint32_t i = (uint32_t)(p - lo - 1) >> 1;
                bits[i >> 5] &= ~(1u << (i & 31));
            }
    }
    if (lo == 0 && tid == 0) bits[0] |= 1u; // 1 is not prime
    // 2a. phases of the sieving numbers nwheel..nsmall, one per thread
    for (uint32_t k = A.nwheel + tid; k < A.nsmall; k += SIEVE_THREADS) {
        const uint64_t i0 = first_index(lo, A.sp[k], A.sp_mu[k]);
        phase[k - A.nwheel] = i0 < SEG_BITS ? (uint32_t)i0 : SEG_BITS;
    }
    __syncthreads();
    // 2b. one wave per sieving number
    const uint32_t nmid = A.nsmall - A.nwheel;
    for (uint32_t k = wave; k < nmid; k += SIEVE_WAVES) {
        const uint32_t p = A.sp[A.nwheel + k];
        for (uint32_t i = phase[k] + lane * p; i < SEG_BITS; i += 64 * p) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    // 3. one larger sieving number per lane (the segment's last candidate
    //    below end saturates: a segment near 2^64 must not wrap)
    const uint64_t hi_num = seg_last(lo, SEG_SPAN, A.end);
    for (uint32_t k = A.nsmall + tid; k < A.nsp; k += SIEVE_THREADS) {
        const uint32_t p = A.sp[k];
        if ((uint64_t)p * p > hi_num) break; // ascending: no later number marks here
        uint64_t i0 = first_index(lo, p, A.sp_mu[k]);
        for (uint64_t i = i0; i < SEG_BITS; i += p) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) words[j] = bits[w0 + j];
    __syncthreads(); // the bitmap is free for staging from here on
}

// Workgroup exclusive scan of one value per lane (wave64 shuffles + LDS).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *wave_sums, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += t;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        uint32_t nw = blockDim.x >> 6;
        uint32_t s = lane < nw ? wave_sums[lane] : 0;
        uint32_t si = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t t = __shfl_up(si, off, 64);
            if (lane >= (uint32_t)off) si += t;
        }
        if (lane < nw) wave_sums[lane] = si - s; // exclusive wave offsets
        if (lane == nw - 1) wave_sums[63] = si;  // grand total
    }
    __syncthreads();
    total = wave_sums[63];
    uint32_t r = wave_sums[wave] + incl - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SIEVE_THREADS) sieve_kernel(SieveArgs A)
{
    __shared__ uint32_t bits[SEG_WORDS];
    __shared__ uint32_t phase[SMALL_MAX_COUNT];
    __shared__ uint32_t wave_sums[64];
    __shared__ unsigned long long seg_base;
    const uint32_t seg = blockIdx.x;
    const uint64_t lo = A.lo0 + (uint64_t)seg * SEG_SPAN;
    uint32_t words[WORDS_PER_THREAD];
    sieve_mark(bits, phase, A, lo, words);

    // survivors of [start, end)
    uint64_t ib = A.start > lo ? (A.start - lo) >> 1 : 0;
    uint64_t ie = A.end > lo ? (A.end - lo) >> 1 : 0;
    if (ib > SEG_BITS) ib = SEG_BITS;
    if (ie > SEG_BITS) ie = SEG_BITS;
    const uint32_t w0 = threadIdx.x * WORDS_PER_THREAD;
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) {
        uint32_t v = ~words[j];
        const uint64_t b0 = (uint64_t)(w0 + j) * 32;
        if (b0 + 32 <= ib || b0 >= ie) v = 0;
        else {
            if (ib > b0) v &= ~0u << (ib - b0);
            if (ie < b0 + 32) v &= (1u << (ie - b0)) - 1u;
        }
        words[j] = v;
        cnt += __popc(v);
    }
    uint32_t total;
    const uint32_t off = block_exclusive_scan(cnt, wave_sums, total);
    if (threadIdx.x == 0) A.seg_counts[seg] = total;
    if (A.mode == SIEVE_COUNT) return;

    uint64_t base;
    if (A.mode != SIEVE_DENSE) {
        base = A.seg_off[seg];
    } else {
        if (threadIdx.x == 0) seg_base = total ? atomicAdd(A.total, (unsigned long long)total) : 0ull;
        __syncthreads();
        base = seg_base;
    }
    // rounds of STAGE_ENTRIES primes: stage bit indices in LDS, then write
    // them out from consecutive lanes
    uint32_t *stage = bits;
    for (uint32_t r0 = 0; r0 < total; r0 += STAGE_ENTRIES) {
        const uint32_t r1 = min(total, r0 + STAGE_ENTRIES);
        if (off < r1 && off + cnt > r0) {
            uint32_t pos = off;
#pragma unroll
            for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) {
                uint32_t v = words[j];
                const uint32_t c = __popc(v);
                if (pos + c <= r0 || pos >= r1) { // whole word outside this round
                    pos += c;
                    continue;
                }
                const uint32_t b0 = (w0 + j) * 32;
                while (v) {
                    const uint32_t bit = __ffs(v) - 1;
                    v &= v - 1;
                    if (pos >= r0 && pos < r1) stage[pos - r0] = b0 + bit;
                    ++pos;
                }
            }
        }
        __syncthreads();
        const uint32_t n = r1 - r0;
        for (uint32_t i = threadIdx.x; i < n; i += SIEVE_THREADS) {
            const uint64_t prime = lo + 2 * (uint64_t)stage[i] + 1;
            const uint64_t at = base + r0 + i;
            if (A.mode == SIEVE_ORDERED32) A.out32[at] = (uint32_t)prime;
            else if (at < A.cap) A.out64[at] = prime;
            // the segment's largest prime is its last staged one
            if (A.mode == SIEVE_DENSE && r0 + i == total - 1) atomicMax(A.maxp, (unsigned long long)prime);
        }
        __syncthreads();
    }
}

// floor(2^64 / p) for each sieving number (p >= 3).
__global__ void sieve_mu_kernel(const uint32_t *sp, uint64_t *mu, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mu[i] = UINT64_MAX / sp[i];
}

// Odd numbers 3, 5, ..., 2n+1 (the first-level sieving list).
__global__ void odd_numbers_kernel(uint32_t *out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 2 * i + 3;
}

} // namespace ctd
