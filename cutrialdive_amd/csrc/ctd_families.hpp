// ctd_families.hpp -- per-family residue state for one prime per lane.
//
// The reference keeps S(n) mod d and steps it with each family's
// next_value_mod (number_sequences/*.hpp).  Here each family keeps the state
// that makes its step cheapest and says when S(n) = 0 mod p:
//
//   family            state at index n              step n -> n+1            S(n) = 0 mod p
//   mersenne          x = 2^n                        x = 2x                   x == 1
//   proth / riesel    x = k 2^n                      x = 2x                   x == p-1 / x == 1
//   fixed-n (k index) x = S(k)                       x += 2^n                 x == 0
//   fibonacci, lucas  (a, b) = (S(n-1), S(n))         (b, a+b)                 b == 0
//   tribonacci        (a, b, c) = (T(n),T(n+1),T(n+2)) (b, c, a+b+c)          a == 0
//   smarandache       r = S(n) (Montgomery form)      r = r b^digits(n+1)+n+1  r == 0
//   test sequence     x = n                           x = x+1                  x == 0
//
// init() computes the state at any index (jump-ahead), so an index range can
// be split into chunks that lanes process independently.  The values are the
// exact residues, hence the hits equal the reference's zero residues, with one
// deliberate exception kept for bit-exactness: fibonacci.hpp:68-71 starts
// index 0 from {F(-1), F(0)} = {0, 0} (lucas.hpp:54-66 inherits it), which
// makes every later residue 0; a job starting at n0 = 0 reproduces that.
#pragma once

#include "ctd_arith.hpp"

namespace ctd {

enum Family : uint32_t {
    FAM_MERSENNE = 0,
    FAM_PROTH = 1,
    FAM_RIESEL = 2,
    FAM_SMARANDACHE = 3,
    FAM_FIBONACCI = 4,
    FAM_LUCAS = 5,
    FAM_PROTH_FIXED_N = 6,
    FAM_RIESEL_FIXED_N = 7,
    FAM_TRIBONACCI = 8,
    FAM_PEANO = 255
};

// JobArgs::flags
constexpr uint32_t JOB_SM_NARROW = 1; // smarandache 32-bit-quotient step allowed (see ctd_tf_fast.hpp)
constexpr uint32_t JOB_FP64 = 2;      // every prime below 2^48: the FP64 forms are allowed (ctd_tf_fp.hpp)
constexpr uint32_t JOB_SM_FULL_STEP = 4; // smarandache FP64 form without its short step (ctd_tf_fp.hpp)
constexpr uint32_t JOB_STEPPED = 16;  // fixed-n families: step every index (no closed form, ctd_tf_fast.hpp)
constexpr uint32_t JOB_SM_NO_SHORT = 32; // smarandache FP64 form: skip the short step (exact-product or general)
constexpr uint32_t JOB_SM_NO_SHIFT = 64; // smarandache FP64 form: skip the shifted step (short, exact-product or general)
constexpr uint32_t JOB_SM_NO_PAIR = 128; // smarandache: one prime per lane (no tf_sm_pair_kernel)

// Job constants visible to every lane.
struct JobArgs {
    uint32_t fam;
    uint32_t maxf;      // max_factors_per_number
    uint64_t param;     // k, fixed n, or base
    uint64_t n0, n1;    // index range
    uint64_t chunk;     // indices per work item
    uint32_t nchunks;
    uint32_t flags;     // JOB_* kernel selections made by the host
    uint32_t *counts;   // [n1-n0]
    uint64_t *factors;  // [(n1-n0) * maxf]
    uint64_t *final_res; // test hook: S(ne-1) mod p per prime index, or null
};

// ---- S(n) mod 2 (prime 2 path) -------------------------------------------
// Closed forms of value_mod_2 for each family.  The reference propagates
// them with next_value_mod_2 (trial_factoring.cuh:139-154); for every
// built-in family that propagation reproduces value_mod_2(n) at each n
// (mersenne.hpp:70-107, riesel_proth.hpp:76-88, riesel_proth_fixed_n.hpp:119-140,
// smarandache.hpp:85-112, fibonacci.hpp:50-61, lucas.hpp:41-51,
// tribonacci.hpp:105-115), so each index can be evaluated on its own.
__device__ __forceinline__ uint32_t value_mod_2(uint32_t fam, uint64_t param, uint64_t n)
{
    switch (fam) {
    case FAM_MERSENNE: return n > 0;
    case FAM_PROTH:
    case FAM_RIESEL: return n > 0 ? 1u : (uint32_t)((param & 1) ^ 1);
    case FAM_PROTH_FIXED_N:
    case FAM_RIESEL_FIXED_N: return 1;
    case FAM_SMARANDACHE:
        if (!(param & 1)) return (uint32_t)(n & 1);
        return (n & 3) && ((n & 3) != 3);
    case FAM_FIBONACCI:
    case FAM_LUCAS: return n % 3 == 0 ? 0u : 1u;
    case FAM_TRIBONACCI: return (uint32_t)((n & 3) >> 1);
    case FAM_PEANO: return (uint32_t)(n & 1);
    }
    return 1;
}

// ---- jump-ahead helpers (Montgomery form) --------------------------------

// (F(m), F(m+1)) by fast doubling.
__device__ inline void fib_pair(const Mont &M, uint64_t m, uint64_t &fa, uint64_t &fb)
{
    const uint64_t p = M.p;
    uint64_t a = 0, b = M.one;
    if (m) {
        for (int bit = 63 - __clzll(m); bit >= 0; --bit) {
            uint64_t c = M.mul(a, sub_mod(dbl_mod(b, p), a, p));   // F(2k)
            uint64_t d = add_mod(M.mul(a, a), M.mul(b, b), p);      // F(2k+1)
            if ((m >> bit) & 1) {
                a = d;
                b = add_mod(c, d, p);
            } else {
                a = c;
                b = d;
            }
        }
    }
    fa = a;
    fb = b;
}

// (T(n), T(n+1), T(n+2)) by the doubling identities of OEIS A000073
// (the ones tribonacci.hpp:64-102 also uses).
__device__ inline void trib_triple(const Mont &M, uint64_t n, uint64_t &ta, uint64_t &tb, uint64_t &tc)
{
    const uint64_t p = M.p;
    uint64_t a = 0, b = 0, c = M.one;
    if (n) {
        for (int bit = 63 - __clzll(n); bit >= 0; --bit) {
            uint64_t aa = M.mul(a, a), bb = M.mul(b, b), cc = M.mul(c, c);
            uint64_t ab2 = dbl_mod(M.mul(a, b), p), ac2 = dbl_mod(M.mul(a, c), p);
            uint64_t bc2 = dbl_mod(M.mul(b, c), p);
            uint64_t t0 = sub_mod(sub_mod(add_mod(ac2, bb, p), ab2, p), aa, p); // T(2k)
            uint64_t t1 = sub_mod(add_mod(aa, bc2, p), bb, p);                  // T(2k+1)
            uint64_t t2 = add_mod(add_mod(ab2, cc, p), bb, p);                  // T(2k+2)
            if ((n >> bit) & 1) {
                a = t1;
                b = t2;
                c = add_mod(add_mod(t0, t1, p), t2, p);
            } else {
                a = t0;
                b = t1;
                c = t2;
            }
        }
    }
    ta = a;
    tb = b;
    tc = c;
}

__device__ __forceinline__ uint64_t sat_mul(uint64_t a, uint64_t b)
{
    return (b && a > UINT64_MAX / b) ? UINT64_MAX : a * b;
}

// Sm_b(n) mod p in Montgomery form, in O(log n) multiplies per digit length:
// appending t consecutive d-digit numbers m0..m0+t-1 maps S to
//   B^t S + m0 U_t + V_t,  B = b^d, U_t = sum_{j<t} B^j, V_t = sum_{i<t} i B^(t-1-i),
// and (B^t, U_t, V_t) follow by doubling on t.  Replaces the host big-integer
// S(n0) and its limb reduction (trial_factoring.cuh:34-47,197-202).
__device__ inline uint64_t sm_value(const Mont &M, uint64_t b, uint64_t n)
{
    const uint64_t p = M.p, one = M.one;
    if (n == 0) return 0;
    uint64_t bM = M.form_of(b);
    uint64_t S = 0;
    uint64_t m0 = 1, m0M = one, BM = bM;
    for (;;) {
        uint64_t m1 = sat_mul(m0, b);
        uint64_t last = n < m1 - 1 ? n : m1 - 1;
        uint64_t t = last - m0 + 1;
        uint64_t P = one, U = 0, V = 0, tM = 0;
        for (int bit = 63 - __clzll(t); bit >= 0; --bit) {
            V = add_mod(add_mod(M.mul(P, V), M.mul(tM, U), p), V, p);
            U = M.mul(U, add_mod(P, one, p));
            P = M.mul(P, P);
            tM = dbl_mod(tM, p);
            if ((t >> bit) & 1) {
                V = add_mod(M.mul(BM, V), tM, p);
                U = add_mod(M.mul(U, BM), one, p);
                P = M.mul(P, BM);
                tM = add_mod(tM, one, p);
            }
        }
        S = add_mod(add_mod(M.mul(P, S), M.mul(m0M, U), p), V, p);
        if (last == n || m1 == UINT64_MAX) break;
        m0 = m1;
        m0M = M.mul(m0M, bM);
        BM = M.mul(BM, bM);
    }
    return S;
}

// digits of m >= 1 in base b, and b^digits(m) (saturating)
__device__ __forceinline__ void digits_pow(uint64_t b, uint64_t m, uint32_t &d, uint64_t &bd)
{
    d = 1;
    bd = b;
    while (bd <= m && bd != UINT64_MAX) {
        bd = sat_mul(bd, b);
        ++d;
    }
}

// ---- generic per-family state (any odd p < 2^64) ---------------------------

struct GenState {
    uint64_t x, y, z; // family-specific (see table above)
    uint64_t c;       // step constant / target
    uint64_t m, nxt;  // smarandache: multiplier (Montgomery) and next digit boundary
    uint64_t nM, bM;  // smarandache: index and base in Montgomery form
};

template <uint32_t FAM>
struct Gen {
    // State at index n (the first index of a work item).
    __device__ static void init(GenState &s, const Mont &M, const JobArgs &J, uint64_t n)
    {
        const uint64_t p = M.p;
        if constexpr (FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL) {
            uint64_t x = M.pow2(n);
            s.x = (FAM == FAM_MERSENNE) ? M.to_plain(x) : M.mul(J.param % p, x);
            s.c = (FAM == FAM_PROTH) ? p - 1 : 1;
        } else if constexpr (FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N) {
            uint64_t cM = M.pow2(J.param);
            s.c = M.to_plain(cM);
            uint64_t x = M.mul(n % p, cM);
            if constexpr (FAM == FAM_PROTH_FIXED_N)
                s.x = (x == p - 1) ? 0 : x + 1;
            else
                s.x = x ? x - 1 : p - 1;
        } else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) {
            if (J.n0 == 0) { // fibonacci.hpp:68-71 quirk: state {0, 0} forever
                s.x = s.y = 0;
                return;
            }
            uint64_t fa, fb;
            fib_pair(M, n - 1, fa, fb);
            fa = M.to_plain(fa);
            fb = M.to_plain(fb);
            if constexpr (FAM == FAM_FIBONACCI) {
                s.x = fa;
                s.y = fb;
            } else { // L(n-1) = 2F(n) - F(n-1), L(n) = F(n) + 2F(n-1)
                s.x = sub_mod(dbl_mod(fb, p), fa, p);
                s.y = add_mod(fb, dbl_mod(fa, p), p);
            }
        } else if constexpr (FAM == FAM_TRIBONACCI) {
            uint64_t a, b, c;
            trib_triple(M, n, a, b, c);
            s.x = M.to_plain(a);
            s.y = M.to_plain(b);
            s.z = M.to_plain(c);
        } else if constexpr (FAM == FAM_SMARANDACHE) {
            const uint64_t b = J.param;
            s.z = b;
            s.x = sm_value(M, b, n);
            s.bM = M.form_of(b);
            s.nM = M.form_of(n % p);
            uint32_t d;
            uint64_t bd;
            digits_pow(b, n + 1, d, bd);
            s.m = M.pow(s.bM, d);
            s.nxt = bd; // first value with more digits than n+1
        } else { // test sequence S(n) = n
            s.x = n % p;
        }
    }

    // Plain residue S(n) mod p of the state at index n (test hook).
    __device__ static uint64_t plain(const GenState &s, const Mont &M)
    {
        const uint64_t p = M.p;
        if constexpr (FAM == FAM_MERSENNE || FAM == FAM_RIESEL) return s.x ? s.x - 1 : p - 1;
        else if constexpr (FAM == FAM_PROTH) return s.x == p - 1 ? 0 : s.x + 1;
        else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) return s.y;
        else if constexpr (FAM == FAM_SMARANDACHE) return M.to_plain(s.x);
        else return s.x;
    }

    __device__ static __forceinline__ bool zero(const GenState &s, const Mont &M)
    {
        if constexpr (FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL) return s.x == s.c;
        else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) return s.y == 0;
        else return s.x == 0;
    }

    // Advance from index n to n+1.
    __device__ static __forceinline__ void step(GenState &s, const Mont &M, uint64_t n)
    {
        const uint64_t p = M.p;
        if constexpr (FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL) {
            s.x = dbl_mod(s.x, p);
        } else if constexpr (FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N) {
            s.x = add_mod(s.x, s.c, p);
        } else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) {
            uint64_t t = add_mod(s.x, s.y, p);
            s.x = s.y;
            s.y = t;
        } else if constexpr (FAM == FAM_TRIBONACCI) {
            uint64_t t = add_mod(add_mod(s.x, s.y, p), s.z, p);
            s.x = s.y;
            s.y = s.z;
            s.z = t;
        } else if constexpr (FAM == FAM_SMARANDACHE) {
            s.nM = add_mod(s.nM, M.one, p);                    // (n+1)_M
            s.x = add_mod(M.mul(s.x, s.m), s.nM, p);           // S(n+1)
            if (n + 2 == s.nxt) {                              // n+2 has one more digit
                s.m = M.mul(s.m, s.bM);
                s.nxt = sat_mul(s.nxt, s.z);                   // z holds the base
            }
        } else {
            s.x = (s.x + 1 == p) ? 0 : s.x + 1;
        }
    }
};

} // namespace ctd
