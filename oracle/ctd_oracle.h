/*
 * ctd_oracle.h -- CPU restatement of cutrialdive's trial-factoring path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the HIP engine's results
 * are compared against, and the CPU baseline timed beside it.  It is never
 * linked into, loaded by or called from the product path (libctd_hip.so, the
 * cutrialdive CLI, the batched engine or the serving loop).
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The arithmetic is plain C with gcc's unsigned __int128.
 */
#ifndef CTD_ORACLE_H
#define CTD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sequence ids: the persisted order of num_seq_id (src/num_seq_id.hpp:16-27),
 * plus one test-only sequence, S(n) = n (tests/peano.hpp:127-137). */
enum {
    CTDO_MERSENNE = 0,
    CTDO_PROTH = 1,
    CTDO_RIESEL = 2,
    CTDO_SMARANDACHE = 3,
    CTDO_FIBONACCI = 4,
    CTDO_LUCAS = 5,
    CTDO_PROTH_FIXED_N = 6,
    CTDO_RIESEL_FIXED_N = 7,
    CTDO_TRIBONACCI = 8,
    CTDO_PEANO = 255
};

/* ---- modular arithmetic (src/modular_arithmetic_detail.hpp) ---------------- */
uint64_t ctdo_mu64(uint64_t p);
void ctdo_mu128(uint64_t p, uint64_t out_lo_hi[2]);
void ctdo_mul128hi(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
uint64_t ctdo_mod(uint64_t a, uint64_t d, uint64_t mu64);
uint64_t ctdo_mulmod(uint64_t a, uint64_t b, uint64_t d);
uint64_t ctdo_modpow(uint64_t a, uint64_t e, uint64_t d);
uint64_t ctdo_reciprocal(uint64_t d);
uint64_t ctdo_mod2by1(uint64_t u1, uint64_t u0, uint64_t d, uint64_t rnd);
uint64_t ctdo_modnby1(const uint64_t *limbs, uint64_t nlen, uint64_t d);
uint64_t ctdo_digits_in_base(uint64_t base, uint64_t n);

/* ---- prime supply (src/siever.hpp:179-237, multi-thread semantics) --------- */
/* Writes the primes of [start, end) in ascending order to out (at most cap of
 * them) and returns how many primes the range holds. */
uint64_t ctdo_sieve(uint64_t start, uint64_t end, uint64_t *out, uint64_t cap);
/* Number of primes in [start, end). */
uint64_t ctdo_prime_count(uint64_t start, uint64_t end);

/* ---- families (src/number_sequences/) ------------------------------------- */
/* S(n) mod 2 via value_mod_2 (number_sequence_helpers.hpp:14-24). */
uint64_t ctdo_value_mod_2(uint32_t seq, uint64_t param, uint64_t n);
/* S(n+1) mod 2 from r = S(n) mod 2 (number_sequence_helpers.hpp:26-37). */
uint64_t ctdo_next_value_mod_2(uint32_t seq, uint64_t param, uint64_t r, uint64_t n);
/* Residues S(n0) mod d ... S(n0+count-1) mod d, computed the way the CPU engine
 * does (trial_factoring_impl.hpp:199-246): init once, then propagate.  d odd. */
int ctdo_residues(uint32_t seq, uint64_t param, uint64_t n0, uint64_t count,
                  uint64_t d, uint64_t *out);
/* For each odd modulus d[i]: out[i] = S(n0+count-1) mod d[i], propagated from
 * n0 exactly as ctdo_residues does (the start value is built once). */
int ctdo_residues_last(uint32_t seq, uint64_t param, uint64_t n0, uint64_t count,
                       const uint64_t *d, uint64_t nd, uint64_t *out);
/* One propagation step S(n+1) mod d from r = S(n) mod d for the stateless
 * families (mersenne, proth, riesel, smarandache, test sequence): the
 * next_value_mod[_mu] functions restated below. */
uint64_t ctdo_next(uint32_t seq, uint64_t param, uint64_t r, uint64_t n, uint64_t d);
/* Family state at index n modulo d (make_state): fibonacci/lucas write
 * {previous, current}, tribonacci {T(n), T(n+1), T(n+2)}, fixed-n {2^n mod d}.
 * Returns the number of words written (0 for stateless families). */
int ctdo_state(uint32_t seq, uint64_t param, uint64_t n, uint64_t d, uint64_t out[3]);
/* Smarandache base b: S(n) as little-endian 64-bit limbs (the HgInt value of
 * smarandache.cpp:129-158).  Returns the limb count; writes at most cap. */
uint64_t ctdo_smarandache_value(uint64_t base, uint64_t n, uint64_t *limbs, uint64_t cap);

/* ---- the engine (src/trial_factoring_impl.hpp:193-366) --------------------- */
typedef struct ctdo_job {
    uint32_t seq;          /* sequence id */
    uint32_t max_factors;  /* max_factors_per_number */
    uint64_t param;        /* k (proth/riesel), n (fixed-n), base (smarandache) */
    uint64_t n0, n1;       /* index range [n0, n1) */
    uint64_t f0, f1;       /* prime range [f0, f1) */
} ctdo_job;

/* Trial-factors job with `threads` workers.  counts[n1-n0] and
 * factors[(n1-n0)*max_factors] use the factors_buffer layout
 * (factors_buffer.hpp:77-86) and are accumulated into (pass zeroed arrays for
 * a fresh run).  Factors of each index are left unsorted, as the buffer holds
 * them.  Returns 0, or -1 on invalid input. */
int ctdo_trial_factor(const ctdo_job *job, uint32_t threads,
                      uint32_t *counts, uint64_t *factors);

/* Primes of [f0, f1) processed, for rate accounting (2 included if in range). */
uint64_t ctdo_job_primes(const ctdo_job *job);

#ifdef __cplusplus
}
#endif
#endif
