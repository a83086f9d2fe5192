// ctd_sieve.hpp -- segmented sieve of Eratosthenes in LDS (gfx950).
//
// Replaces the host prime supply: siever::sieve over primesieve
// (siever.hpp:179-237) and the per-segment H2D copy of
// device_prime_data::copy_from_host (prime_data.hpp:165-196).
//
// One workgroup sieves one segment of SEG_SPAN integers [lo, lo+SEG_SPAN),
// lo even.  Bit i of the LDS bitmap stands for the odd number lo + 2i + 1.
// Sieving primes below SMALL_LIMIT are marked cooperatively by the whole
// workgroup (many multiples each), larger ones one per lane (at most
// SEG_BITS / p multiples).  Survivors are compacted in ascending order with a
// workgroup prefix sum of per-lane popcounts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctd {

constexpr uint32_t SEG_BITS = 1u << 19;            // odd numbers per segment
constexpr uint64_t SEG_SPAN = 2ull * SEG_BITS;      // integers per segment
constexpr uint32_t SEG_WORDS = SEG_BITS / 32;       // 16384 words = 64 KiB of LDS
constexpr uint32_t SIEVE_THREADS = 1024;
constexpr uint32_t WORDS_PER_THREAD = SEG_WORDS / SIEVE_THREADS;
constexpr uint32_t SMALL_LIMIT = 1u << 13;          // cooperative marking below this

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
    uint32_t pad;
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

__device__ __forceinline__ void sieve_mark(uint32_t *bits, const SieveArgs &A, uint64_t lo)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t w = tid; w < SEG_WORDS; w += SIEVE_THREADS) bits[w] = 0;
    __syncthreads();
    // cooperative marking of small sieving numbers
    for (uint32_t k = 0; k < A.nsmall; ++k) {
        uint32_t p = A.sp[k];
        uint64_t i0 = first_index(lo, p, A.sp_mu[k]);
        for (uint64_t i = i0 + (uint64_t)tid * p; i < SEG_BITS; i += (uint64_t)SIEVE_THREADS * p)
            atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    // one larger sieving number per lane
    const uint64_t hi_num = lo + SEG_SPAN - 1;
    for (uint32_t k = A.nsmall + tid; k < A.nsp; k += SIEVE_THREADS) {
        uint32_t p = A.sp[k];
        if ((uint64_t)p * p > hi_num) break; // ascending: no later number marks here
        uint64_t i0 = first_index(lo, p, A.sp_mu[k]);
        for (uint64_t i = i0; i < SEG_BITS; i += p) atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
    if (lo == 0 && tid == 0) bits[0] |= 1u; // 1 is not prime
    __syncthreads();
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
    __shared__ uint32_t wave_sums[64];
    __shared__ unsigned long long seg_base;
    const uint32_t seg = blockIdx.x;
    const uint64_t lo = A.lo0 + (uint64_t)seg * SEG_SPAN;
    sieve_mark(bits, A, lo);

    // valid bit range for [start, end)
    uint64_t ib = A.start > lo ? (A.start - lo) >> 1 : 0;
    uint64_t ie = A.end > lo ? (A.end - lo) >> 1 : 0;
    if (ib > SEG_BITS) ib = SEG_BITS;
    if (ie > SEG_BITS) ie = SEG_BITS;

    const uint32_t w0 = threadIdx.x * WORDS_PER_THREAD;
    uint32_t words[WORDS_PER_THREAD];
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) {
        uint32_t w = w0 + j;
        uint32_t v = ~bits[w];
        uint64_t b0 = (uint64_t)w * 32;
        if (b0 + 32 <= ib || b0 >= ie) v = 0;
        else {
            if (ib > b0) v &= ~0u << (ib - b0);
            if (ie < b0 + 32) v &= (1u << (ie - b0)) - 1u;
        }
        words[j] = v;
        cnt += __popc(v);
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(cnt, wave_sums, total);
    if (threadIdx.x == 0) A.seg_counts[seg] = total;
    if (A.mode == SIEVE_COUNT) return;

    uint64_t base;
    if (A.mode != SIEVE_DENSE) {
        base = A.seg_off[seg];
    } else {
        if (threadIdx.x == 0) {
            seg_base = total ? atomicAdd(A.total, (unsigned long long)total) : 0ull;
        }
        __syncthreads();
        base = seg_base;
    }
    uint64_t pos = base + off;
    uint64_t last = 0;
#pragma unroll
    for (uint32_t j = 0; j < WORDS_PER_THREAD; ++j) {
        uint32_t v = words[j];
        uint64_t b0 = (uint64_t)(w0 + j) * 32;
        while (v) {
            uint32_t bit = __ffs(v) - 1;
            v &= v - 1;
            uint64_t prime = lo + 2 * (b0 + bit) + 1;
            if (A.mode == SIEVE_ORDERED32) {
                A.out32[pos] = (uint32_t)prime;
            } else if (pos < A.cap) {
                A.out64[pos] = prime;
            }
            last = prime;
            ++pos;
        }
    }
    if (A.mode == SIEVE_DENSE && last) atomicMax(A.maxp, (unsigned long long)last);
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
