// ctd_tf_fast.hpp -- narrow per-family index loops for primes below 2^63.
//
// Set-up (powers, doubling formulas, jump-ahead) is the Montgomery code of
// ctd_families.hpp; the per-index step then uses the cheapest exact form for
// the family.  The index loop is wave-uniform (scalar counters, one scalar
// branch per UNROLL indices); all per-index work is VALU:
//
//   doubling, 32-bit residues : shift, subtract, min, compare        (4 ops)
//   doubling, 64-bit residues : shift, sub/borrow x2, select x2, cmp (6 ops)
//   smarandache, p >= 2^32    : 32-bit-quotient Barrett on r*b^d + n+1 with
//                               full-rate 24-bit multiplies, lazy [0,2p) (~12 ops)
#pragma once

#include "ctd_families.hpp"

namespace ctd {

// One atomic per wave for the lanes whose residue is zero at index idx
// (device_factors_buffer.hpp:119-129 does one atomic per hit).  All active
// lanes of the wave are at the same index when this is called.
__device__ __forceinline__ void record_hits_mask(const JobArgs &J, uint64_t idx, uint64_t p, bool hit,
                                                 uint64_t mask)
{
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

__device__ __forceinline__ void record_hits(const JobArgs &J, uint64_t idx, uint64_t p, bool hit)
{
    const uint64_t mask = __ballot(hit);
    if (mask) record_hits_mask(J, idx, p, hit, mask);
}

constexpr uint64_t FAST_PMAX = 1ull << 63;
constexpr int UNROLL = 4;

// True when every lane of the wave holds a prime below 2^31.
__device__ __forceinline__ bool wave_w32(uint64_t p) { return __ballot(p >= (1ull << 31)) == 0; }

// u >= q ? u - q : u for u < 2q, q < 2^63, u given as 32-bit halves.  The
// borrow of the 64-bit subtraction selects directly (the compiler would
// re-derive it with an extra 64-bit compare).  Two wait states separate each
// VCC write from its VALU reader, as gfx950 requires.
__device__ __forceinline__ uint64_t sub_if_ge2(uint32_t ul, uint32_t uh, uint64_t q)
{
    const uint32_t ql = (uint32_t)q, qh = (uint32_t)(q >> 32);
    uint32_t rl, rh;
    asm("v_sub_co_u32 %0, vcc, %2, %4\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32 %1, vcc, %3, %5, vcc\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32 %0, %0, %2, vcc\n\t"
        "v_cndmask_b32 %1, %1, %3, vcc"
        : "=&v"(rl), "=&v"(rh)
        : "v"(ul), "v"(uh), "v"(ql), "v"(qh)
        : "vcc");
    return ((uint64_t)rh << 32) | rl;
}
__device__ __forceinline__ uint64_t sub_if_ge(uint64_t u, uint64_t q)
{
    return sub_if_ge2((uint32_t)u, (uint32_t)(u >> 32), q);
}

// 2x mod p for x < p < 2^63
__device__ __forceinline__ uint64_t dbl_fast(uint64_t x, uint64_t p) { return sub_if_ge(x << 1, p); }
// a + b mod p for a, b < p < 2^63
__device__ __forceinline__ uint64_t add_fast(uint64_t a, uint64_t b, uint64_t p) { return sub_if_ge(a + b, p); }

// v_mad_u32_u24 / v_mad_i32_i24: 24-bit multiply-add, full rate
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ uint32_t mad_i24(uint32_t a, int32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ uint32_t add32(uint32_t a, uint32_t b, uint32_t q)
{
    uint32_t t = a + b;
    return min(t, t - q);
}

// idx + k as an opaque scalar: the hit path computes its addresses from it,
// so the compiler keeps no per-group pointer induction variables (scalar
// work that would compete with the VALU stream).
__device__ __forceinline__ uint64_t hit_base(uint64_t idx, uint32_t k)
{
    uint64_t at = idx + k;
    asm volatile("" : "+s"(at));
    return at;
}

// Runs `steps` steps with one wave-wide branch per UNROLL indices.  STEP4
// sets the bools h0..h3 of the 4 unrolled indices; STEP1 does one step whose
// hit is HIT1.  idx is the index (relative to n0) of the state before the
// loop; the counter is 32-bit (work items hold fewer than 2^31 indices) so
// the loop control stays three scalar instructions.
#define CTD_UNROLLED_LOOP(STEP4, STEP1, HIT1)                                                    \
    {                                                                                              \
        const uint32_t steps32 = (uint32_t)steps;                                                 \
        uint32_t k = 0;                                                                            \
        for (; k + UNROLL <= steps32; k += UNROLL) {                                               \
            bool h0, h1, h2, h3;                                                                   \
            STEP4;                                                                                 \
            const uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2),               \
                           m3 = __ballot(h3);                                                      \
            if (m0 | m1 | m2 | m3) {                                                               \
                const uint64_t at = hit_base(idx, k);                                              \
                if (m0) record_hits_mask(J, at + 1, p, h0, m0);                                    \
                if (m1) record_hits_mask(J, at + 2, p, h1, m1);                                    \
                if (m2) record_hits_mask(J, at + 3, p, h2, m2);                                    \
                if (m3) record_hits_mask(J, at + 4, p, h3, m3);                                    \
            }                                                                                      \
        }                                                                                          \
        idx += k;                                                                                  \
        for (; k < steps32; ++k) {                                                                 \
            STEP1;                                                                                 \
            ++idx;                                                                                 \
            record_hits(J, idx, p, HIT1);                                                          \
        }                                                                                          \
    }

// The same with the step computing the wave masks m0..m3 itself (a hit that
// is the OR of two compares stays a pair of scalar masks).
#define CTD_UNROLLED_LOOP_MASKS(STEP4, STEP1, MASK1)                                             \
    {                                                                                              \
        const uint32_t steps32 = (uint32_t)steps;                                                 \
        uint32_t k = 0;                                                                            \
        for (; k + UNROLL <= steps32; k += UNROLL) {                                               \
            uint64_t m0, m1, m2, m3;                                                               \
            STEP4;                                                                                 \
            if (m0 | m1 | m2 | m3) {                                                               \
                const uint32_t lane = __lane_id();                                                 \
                const uint64_t at = hit_base(idx, k);                                              \
                if (m0) record_hits_mask(J, at + 1, p, (m0 >> lane) & 1, m0);                      \
                if (m1) record_hits_mask(J, at + 2, p, (m1 >> lane) & 1, m1);                      \
                if (m2) record_hits_mask(J, at + 3, p, (m2 >> lane) & 1, m2);                      \
                if (m3) record_hits_mask(J, at + 4, p, (m3 >> lane) & 1, m3);                      \
            }                                                                                      \
        }                                                                                          \
        idx += k;                                                                                  \
        for (; k < steps32; ++k) {                                                                 \
            STEP1;                                                                                 \
            ++idx;                                                                                 \
            const uint64_t mk = MASK1;                                                             \
            if (mk) record_hits_mask(J, idx, p, (mk >> __lane_id()) & 1, mk);                      \
        }                                                                                          \
    }

// Eight indices per scalar branch: for the 4-5 instruction steps the mask
// combining and loop control would otherwise keep the CU's scalar unit as
// busy as its VALUs.  STEP8 sets the masks mk[0..7].
#define CTD_UNROLLED_LOOP_MASKS8(STEP8, STEP1, MASK1)                                            \
    {                                                                                              \
        const uint32_t steps32 = (uint32_t)steps;                                                 \
        uint32_t k = 0;                                                                            \
        for (; k + 8 <= steps32; k += 8) {                                                         \
            uint64_t mk[8];                                                                        \
            STEP8;                                                                                 \
            if ((mk[0] | mk[1] | mk[2] | mk[3]) | (mk[4] | mk[5] | mk[6] | mk[7])) {               \
                const uint32_t lane = __lane_id();                                                 \
                const uint64_t at = hit_base(idx, k);                                              \
                for (int j = 0; j < 8; ++j)                                                        \
                    if (mk[j]) record_hits_mask(J, at + 1 + j, p, (mk[j] >> lane) & 1, mk[j]);     \
            }                                                                                      \
        }                                                                                          \
        idx += k;                                                                                  \
        for (; k < steps32; ++k) {                                                                 \
            STEP1;                                                                                 \
            ++idx;                                                                                 \
            const uint64_t m1 = MASK1;                                                             \
            if (m1) record_hits_mask(J, idx, p, (m1 >> __lane_id()) & 1, m1);                      \
        }                                                                                          \
    }

// B indices per scalar branch (B = 12 or 16), for the stride forms whose
// VALU work per index is one compare: B - 1 mask ORs and 3 loop instructions
// on the scalar unit against B + 3 to B + 6 VALU instructions.  STEPB sets the
// masks mk[0..B-1].
#define CTD_UNROLLED_LOOP_MASKSB(B, STEPB, STEP1, MASK1)                                          \
    {                                                                                              \
        const uint32_t steps32 = (uint32_t)steps;                                                 \
        uint32_t k = 0;                                                                            \
        for (; k + (B) <= steps32; k += (B)) {                                                     \
            uint64_t mk[B];                                                                        \
            STEPB;                                                                                 \
            uint64_t any = 0;                                                                      \
            _Pragma("unroll") for (int j = 0; j < (B); ++j) any |= mk[j];                         \
            if (any) {                                                                             \
                const uint32_t lane = __lane_id();                                                 \
                const uint64_t at = hit_base(idx, k);                                              \
                for (int j = 0; j < (B); ++j)                                                      \
                    if (mk[j]) record_hits_mask(J, at + 1 + j, p, (mk[j] >> lane) & 1, mk[j]);     \
            }                                                                                      \
        }                                                                                          \
        idx += k;                                                                                  \
        for (; k < steps32; ++k) {                                                                 \
            STEP1;                                                                                 \
            ++idx;                                                                                 \
            const uint64_t m1 = MASK1;                                                             \
            if (m1) record_hits_mask(J, idx, p, (m1 >> __lane_id()) & 1, m1);                      \
        }                                                                                          \
    }
#define CTD_UNROLLED_LOOP_MASKS16(STEP16, STEP1, MASK1) CTD_UNROLLED_LOOP_MASKSB(16, STEP16, STEP1, MASK1)

} // namespace ctd

#include "ctd_tf_fp.hpp"

namespace ctd {

// The FP64 form (ctd_tf_fp.hpp) takes a wave when the host allows it
// (JOB_FP64: every prime of the batch below FP_PMAX) and it is the cheaper
// form for the wave: the 32-bit integer steps stay cheaper for
// Fibonacci/Lucas when every prime is below 2^31, unless the work item is
// long enough for the theta stride form (the stride forms of the doubling and
// fixed-n families beat them at every size); Smarandache needs every prime
// above 8 n1 and the narrow multiplier bound.
template <uint32_t FAM>
__device__ __forceinline__ bool use_fp(const JobArgs &J, uint64_t p, uint64_t steps)
{
    if (!(J.flags & JOB_FP64) || __ballot(p >= FP_PMAX) != 0) return false;
    if constexpr (FAM == FAM_PEANO) return false;
    else if constexpr (FAM == FAM_TRIBONACCI || FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL ||
                       FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N)
        return true; // the stride forms beat the 4-instruction 32-bit steps
    else if constexpr (FAM == FAM_SMARANDACHE)
        return (J.flags & JOB_SM_NARROW) && __ballot(p <= 8 * J.n1) == 0;
    else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS)
        return !wave_w32(p) || (J.n0 != 0 && steps >= THETA_MIN_STEPS && __ballot(p == 5) == 0);
    else return !wave_w32(p);
}

// One lane's hit at idx (lanes of a wave may be at different indices).
__device__ __forceinline__ void record_hit_lane(const JobArgs &J, uint64_t idx, uint64_t p)
{
    const uint32_t pos = atomicAdd(&J.counts[idx], 1u);
    if (pos < J.maxf) J.factors[idx * J.maxf + pos] = p;
}

// Fixed n: S(k) = k 2^n + 1 (proth) or k 2^n - 1 (riesel) vanishes mod p
// exactly when k = -2^-n or k = 2^-n (mod p), so the hits among the indices
// [ns, ne) are that residue class; one power per prime replaces the ne - ns
// steps S(k+1) = S(k) + 2^n of riesel_proth_fixed_n.hpp:142-151 (the stepped
// forms stay available as JOB_STEPPED).  2^-n is ((p + 1) / 2)^n.
template <uint32_t FAM>
__device__ void run_fixed_n_closed(const JobArgs &J, uint64_t p, uint64_t ns, uint64_t ne)
{
    const uint64_t half = (p >> 1) + 1; // 2^-1 mod p
    uint64_t u, r;                      // 2^-n mod p, ns mod p (both in [0, p))
    if ((J.flags & JOB_FP64) && __ballot(p >= FP_PMAX) == 0 && __ballot(ns >= (1ull << 52)) == 0) {
        Fmod F;
        F.init(p);
        u = F.canon(F.pow((double)half, J.param), p);
        r = F.canon(F.of(ns), p);
    } else {
        Mont M;
        M.init(p);
        u = M.to_plain(M.pow(M.form_of(half), J.param));
        r = ns % p;
    }
    const uint64_t ks = (FAM == FAM_PROTH_FIXED_N) ? (u ? p - u : 0) : u;
    const uint64_t span = ne - ns;
    for (uint64_t d = ks >= r ? ks - r : ks + (p - r); d < span; d += p) {
        record_hit_lane(J, ns + d - J.n0, p);
        if (span - 1 - d < p) break; // the next member is at or past ne (and d + p could wrap)
    }
}

template <uint32_t FAM>
struct Fast {
    __host__ __device__ static bool fits(uint64_t pmax) { return pmax < FAST_PMAX; }

    // ns, ne are wave-uniform; pi is the prime's index in the batch.
    __device__ static void run(const JobArgs &J, uint64_t p, uint64_t ns, uint64_t ne, uint64_t pi)
    {
        // test hook: the plain residue S(ne-1) mod p
        auto emit = [&](uint64_t v) {
            if (J.final_res) J.final_res[pi] = v;
        };
        if constexpr (FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N) {
            if (!(J.flags & JOB_STEPPED)) {
                run_fixed_n_closed<FAM>(J, p, ns, ne);
                if (J.final_res) {
                    Mont M;
                    M.init(p);
                    GenState s;
                    Gen<FAM>::init(s, M, J, ne - 1);
                    emit(Gen<FAM>::plain(s, M));
                }
                return;
            }
        }
        if (use_fp<FAM>(J, p, ne - ns - 1)) {
            emit(run_fp<FAM>(J, p, ns, ne, ns - J.n0, ne - ns - 1));
            return;
        }
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
                CTD_UNROLLED_LOOP(
                    {
                        y = add32(y, y, q); h0 = y == t;
                        y = add32(y, y, q); h1 = y == t;
                        y = add32(y, y, q); h2 = y == t;
                        y = add32(y, y, q); h3 = y == t;
                    },
                    y = add32(y, y, q), y == t)
                x = y;
            } else {
                CTD_UNROLLED_LOOP(
                    {
                        x = dbl_fast(x, p); h0 = x == tgt;
                        x = dbl_fast(x, p); h1 = x == tgt;
                        x = dbl_fast(x, p); h2 = x == tgt;
                        x = dbl_fast(x, p); h3 = x == tgt;
                    },
                    x = dbl_fast(x, p), x == tgt)
            }
            // S = x - 1 (mersenne, riesel) or x + 1 (proth)
            if constexpr (FAM == FAM_PROTH) emit(x == p - 1 ? 0 : x + 1);
            else emit(x ? x - 1 : p - 1);
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
                CTD_UNROLLED_LOOP(
                    {
                        y = add32(y, cc, q); h0 = y == 0;
                        y = add32(y, cc, q); h1 = y == 0;
                        y = add32(y, cc, q); h2 = y == 0;
                        y = add32(y, cc, q); h3 = y == 0;
                    },
                    y = add32(y, cc, q), y == 0)
                x = y;
            } else {
                CTD_UNROLLED_LOOP(
                    {
                        x = add_fast(x, c, p); h0 = x == 0;
                        x = add_fast(x, c, p); h1 = x == 0;
                        x = add_fast(x, c, p); h2 = x == 0;
                        x = add_fast(x, c, p); h3 = x == 0;
                    },
                    x = add_fast(x, c, p), x == 0)
            }
            emit(x);
        } else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) {
            GenState s;
            Gen<FAM>::init(s, M, J, ns);
            record_hits(J, idx, p, s.y == 0);
            if (wave_w32(p)) {
                const uint32_t q = (uint32_t)p;
                uint32_t a = (uint32_t)s.x, b = (uint32_t)s.y;
                CTD_UNROLLED_LOOP(
                    {
                        a = add32(a, b, q); h0 = a == 0;
                        b = add32(a, b, q); h1 = b == 0;
                        a = add32(a, b, q); h2 = a == 0;
                        b = add32(a, b, q); h3 = b == 0;
                    },
                    { uint32_t t = add32(a, b, q); a = b; b = t; }, b == 0)
                emit(b);
            } else {
                uint64_t a = s.x, b = s.y;
                CTD_UNROLLED_LOOP(
                    {
                        a = add_fast(a, b, p); h0 = a == 0;
                        b = add_fast(a, b, p); h1 = b == 0;
                        a = add_fast(a, b, p); h2 = a == 0;
                        b = add_fast(a, b, p); h3 = b == 0;
                    },
                    { uint64_t t = add_fast(a, b, p); a = b; b = t; }, b == 0)
                emit(b);
            }
        } else if constexpr (FAM == FAM_TRIBONACCI) {
            GenState s;
            Gen<FAM>::init(s, M, J, ns);
            record_hits(J, idx, p, s.x == 0);
            // (a, b, c) = (T(n), T(n+1), T(n+2)); the new value replaces the
            // oldest register so the unrolled body needs no moves
            if (wave_w32(p)) {
                const uint32_t q = (uint32_t)p;
                uint32_t a = (uint32_t)s.x, b = (uint32_t)s.y, c = (uint32_t)s.z;
                CTD_UNROLLED_LOOP(
                    {
                        a = add32(add32(a, b, q), c, q); h0 = b == 0;
                        b = add32(add32(b, c, q), a, q); h1 = c == 0;
                        c = add32(add32(c, a, q), b, q); h2 = a == 0;
                        uint32_t t = add32(add32(a, b, q), c, q);
                        a = b; b = c; c = t; h3 = a == 0;
                    },
                    { uint32_t t = add32(add32(a, b, q), c, q); a = b; b = c; c = t; }, a == 0)
                emit(a);
            } else {
                uint64_t a = s.x, b = s.y, c = s.z;
                CTD_UNROLLED_LOOP(
                    {
                        a = add_fast(add_fast(a, b, p), c, p); h0 = b == 0;
                        b = add_fast(add_fast(b, c, p), a, p); h1 = c == 0;
                        c = add_fast(add_fast(c, a, p), b, p); h2 = a == 0;
                        uint64_t t = add_fast(add_fast(a, b, p), c, p);
                        a = b; b = c; c = t; h3 = a == 0;
                    },
                    { uint64_t t = add_fast(add_fast(a, b, p), c, p); a = b; b = c; c = t; }, a == 0)
                emit(a);
            }
        } else if constexpr (FAM == FAM_SMARANDACHE) {
            emit(run_smarandache(J, M, p, ns, ne, idx, steps));
        } else {
            GenState s;
            Gen<FAM>::init(s, M, J, ns);
            record_hits(J, idx, p, Gen<FAM>::zero(s, M));
            for (uint64_t n = ns; n + 1 < ne; ++n) {
                Gen<FAM>::step(s, M, n);
                ++idx;
                record_hits(J, idx, p, Gen<FAM>::zero(s, M));
            }
            emit(Gen<FAM>::plain(s, M));
        }
    }

    // S(n+1) = S(n) b^digits(n+1) + n+1.  With m = b^digits(n+1) < 2^21 and
    // k = bitlen(p) in [33, 54] (checked by the host: JOB_SM_NARROW), the
    // quotient of r*m by p is below 2^23 and is estimated from the top 32
    // bits of r and W = floor(m 2^(k+1) / p):  q = mulhi(r >> (k-31), W)
    // lies in [Q-2, Q] for Q = floor(r m / p), so y = r m + n+1 - q p lies in
    // [0, 4p) and one conditional subtraction of 2p keeps r in [0, 2p).
    // S(n) = 0 mod p  <=>  r in {0, p}.
    // Returns the plain residue S(ne-1) mod p.
    __device__ static uint64_t run_smarandache(const JobArgs &J, const Mont &M, uint64_t p, uint64_t ns,
                                               uint64_t ne, uint64_t idx, uint64_t steps)
    {
        GenState s;
        Gen<FAM_SMARANDACHE>::init(s, M, J, ns);
        record_hits(J, idx, p, s.x == 0);
        const uint64_t b = J.param;
        const bool narrow = (J.flags & JOB_SM_NARROW) && __ballot(p < (1ull << 32)) == 0;
        if (!narrow) {
            for (uint64_t n = ns; n + 1 < ne; ++n) {
                Gen<FAM_SMARANDACHE>::step(s, M, n);
                ++idx;
                record_hits(J, idx, p, s.x == 0);
            }
            return M.to_plain(s.x);
        }
        // wave-uniform digit state for index n = ns: m = b^digits(n+1)
        uint32_t d;
        uint64_t nxt;
        digits_pow(b, ns + 1, d, nxt);
        uint32_t m = 1;
        for (uint32_t i = 0; i < d; ++i) m *= (uint32_t)b;
        const uint32_t k = 64 - __clzll(p);
        const uint32_t sh = k - 31;
        const uint32_t plneg = 0u - (uint32_t)p;               // 2^32 - p_lo (p odd)
        const int32_t nph1 = -(int32_t)((uint32_t)(p >> 32) + 1); // -(p_hi + 1)
        const uint64_t p2 = p << 1;
        uint32_t W = (uint32_t)(((uint64_t)m << (k + 1)) / p);
        uint64_t r = M.to_plain(s.x); // the jump-ahead leaves S(ns) in Montgomery form
        uint64_t c = ns + 1;          // the number appended by the next step (wave-uniform)
        auto step = [&](void) {
            const uint32_t rl = (uint32_t)r, rh = (uint32_t)(r >> 32);
            uint32_t q;
            uint64_t t;
            // q = mulhi(r >> sh, W); t = q (2^32 - p_lo) + rl m + c
            asm("v_alignbit_b32 %1, %3, %2, %4\n\t"
                "v_mul_hi_u32 %1, %1, %5\n\t"
                "v_mad_u64_u32 %0, vcc, %2, %6, %7\n\t"
                "v_mad_u64_u32 %0, vcc, %1, %8, %0"
                : "=&v"(t), "=&v"(q)
                : "v"(rl), "v"(rh), "v"(sh), "v"(W), "v"(m), "s"(c), "v"(plneg)
                : "vcc");
            // high word: + rh m - q (p_hi + 1); then r = y mod 2p
            const uint32_t yh = mad_i24(q, nph1, mad_u24(rh, m, (uint32_t)(t >> 32)));
            r = sub_if_ge2((uint32_t)t, yh, p2);
        };
        auto hits = [&](void) { return __ballot(r == 0) | __ballot(r == p); };
        uint64_t n = ns; // index of the current state
        while (n + 1 < ne) {
            // indices n+1 .. seg_end use the current m (digits change when n+2 == nxt)
            uint64_t seg_last = nxt - 2 < ne - 1 ? nxt - 2 : ne - 2; // last n of a step with this m
            const uint32_t cnt = (uint32_t)(seg_last - n + 1); // below 2^32: n1 <= 2^32
            uint32_t kk = 0;
            for (; kk + UNROLL <= cnt; kk += UNROLL) {
                step(); c++; const uint64_t m0 = hits();
                step(); c++; const uint64_t m1 = hits();
                step(); c++; const uint64_t m2 = hits();
                step(); c++; const uint64_t m3 = hits();
                if (m0 | m1 | m2 | m3) {
                    const uint32_t lane = __lane_id();
                    const uint64_t at = hit_base(idx, kk);
                    if (m0) record_hits_mask(J, at + 1, p, (m0 >> lane) & 1, m0);
                    if (m1) record_hits_mask(J, at + 2, p, (m1 >> lane) & 1, m1);
                    if (m2) record_hits_mask(J, at + 3, p, (m2 >> lane) & 1, m2);
                    if (m3) record_hits_mask(J, at + 4, p, (m3 >> lane) & 1, m3);
                }
            }
            idx += kk;
            for (; kk < cnt; ++kk) {
                step();
                c++;
                ++idx;
                const uint64_t mk = hits();
                if (mk) record_hits_mask(J, idx, p, (mk >> __lane_id()) & 1, mk);
            }
            n = seg_last + 1;
            if (n + 1 < ne) { // the next appended number n+1 has one more digit
                m *= (uint32_t)b;
                nxt = sat_mul(nxt, b);
                W = (uint32_t)(((uint64_t)m << (k + 1)) / p);
            }
        }
        return r >= p ? r - p : r;
    }
};

} // namespace ctd
