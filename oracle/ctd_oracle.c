/*
 * ctd_oracle.c -- CPU restatement of cutrialdive's trial-factoring path.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP engine's results are compared
 * against, and the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Nothing
 * on the product path does.
 *
 * Each function restates the reference behaviour it cites (file:line under
 * the reference's src/).  Where the reference calls a third-party library
 * (primesieve for primes, GMP for the Smarandache start value), the published
 * mathematics is restated instead: an exact segmented sieve of Eratosthenes
 * (with a deterministic Miller-Rabin test for tiny windows near 2^64) and
 * exact schoolbook big-integer concatenation.
 */
#include "ctd_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ======================================================================== */
/* Modular arithmetic: src/modular_arithmetic_detail.hpp                    */
/* ======================================================================== */

/* mu64(p) = floor((2^64-1)/p)   (modular_arithmetic_detail.hpp:249-253) */
uint64_t ctdo_mu64(uint64_t p) { return UINT64_MAX / p; }

static u128 mu128_u(uint64_t p) { return (~(u128)0) / p; }

/* mu128(p) = floor((2^128-1)/p)  (modular_arithmetic_detail.hpp:255-259) */
void ctdo_mu128(uint64_t p, uint64_t out[2])
{
    u128 m = mu128_u(p);
    out[0] = (uint64_t)m;
    out[1] = (uint64_t)(m >> 64);
}

/* High 128 bits of a 128x128 product (modular_arithmetic_detail.hpp:266-289).
 * Schoolbook over 64-bit halves: the middle column is summed in 128 bits so
 * its carry into the high half is exact. */
static u128 mul128hi_u(u128 a, u128 b)
{
    uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64);
    uint64_t b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1;
    u128 p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
    u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    return p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
}

void ctdo_mul128hi(const uint64_t a[2], const uint64_t b[2], uint64_t out[2])
{
    u128 x = ((u128)a[1] << 64) | a[0];
    u128 y = ((u128)b[1] << 64) | b[0];
    u128 r = mul128hi_u(x, y);
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}

/* a mod d by Barrett with mu64, one correction (modular_arithmetic_detail.hpp:335-340) */
uint64_t ctdo_mod(uint64_t a, uint64_t d, uint64_t mu64)
{
    uint64_t q = (uint64_t)(((u128)a * mu64) >> 64);
    uint64_t r = a - q * d;
    return r >= d ? r - d : r;
}

/* Barrett context of one divisor: the mu_both_t of barrett_mu_types.hpp:28-32 */
typedef struct { uint64_t d, mu64; u128 mu128; } bctx;

static bctx bctx_make(uint64_t d)
{
    bctx c;
    c.d = d;
    c.mu64 = ctdo_mu64(d);
    c.mu128 = mu128_u(d);
    return c;
}

/* a*b mod d: mu64 when the product fits 64 bits, mu128 otherwise, one
 * correction (modular_arithmetic_detail.hpp:349-355) */
static uint64_t mulmod_b(uint64_t a, uint64_t b, const bctx *c)
{
    u128 ab = (u128)a * b;
    u128 q = (ab >> 64) ? mul128hi_u(ab, c->mu128) : ((ab * c->mu64) >> 64);
    u128 r = ab - q * c->d;
    return (uint64_t)(r >= c->d ? r - c->d : r);
}

uint64_t ctdo_mulmod(uint64_t a, uint64_t b, uint64_t d)
{
    bctx c = bctx_make(d);
    return mulmod_b(a, b, &c);
}

/* 2a, a+b, a-b mod d (modular_arithmetic_detail.hpp:376-403) */
static uint64_t doublemod(uint64_t a, uint64_t d) { return a < d - a ? a << 1 : a - (d - a); }
static uint64_t addmod(uint64_t a, uint64_t b, uint64_t d) { return a < d - b ? a + b : a - (d - b); }
static uint64_t submod(uint64_t a, uint64_t b, uint64_t d) { return a >= b ? a - b : d - (b - a); }

/* a^e mod d, right-to-left binary (modular_arithmetic_detail.hpp:413-424) */
static uint64_t modpow_b(uint64_t a, uint64_t e, const bctx *c)
{
    uint64_t r = 1;
    uint64_t m = a < c->d ? a : ctdo_mod(a, c->d, c->mu64);
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod_b(r, m, c);
        m = mulmod_b(m, m, c);
    }
    return r;
}

uint64_t ctdo_modpow(uint64_t a, uint64_t e, uint64_t d)
{
    bctx c = bctx_make(d);
    return modpow_b(a, e, &c);
}

/* floor((2^128-1)/d) - 2^64 for 2^63 <= d (modular_arithmetic_detail.hpp:155-177,
 * host branch) */
uint64_t ctdo_reciprocal(uint64_t d)
{
    u128 num = ((u128)(UINT64_MAX - d) << 64) | UINT64_MAX;
    return (uint64_t)(num / d);
}

/* (u1*2^64 + u0) mod d for normalised d, u1 < d, with the reciprocal rnd:
 * Moller-Granlund 2-by-1 remainder ("Improved division by invariant
 * integers", 2011, Algorithm 4), as used at modular_arithmetic_detail.hpp:181-211 */
uint64_t ctdo_mod2by1(uint64_t u1, uint64_t u0, uint64_t d, uint64_t rnd)
{
    u128 q = (u128)rnd * u1 + (((u128)u1 << 64) | u0);
    uint64_t q1 = (uint64_t)(q >> 64) + 1, q0 = (uint64_t)q;
    uint64_t r = u0 - q1 * d;
    if (r > q0) r += d;
    if (r >= d) r -= d;
    return r;
}

/* N mod d for N = sum limbs[i] 2^(64 i): reduce modulo the normalised divisor
 * d 2^s with 2-by-1 steps, then modulo d (modular_arithmetic_detail.hpp:213-232) */
uint64_t ctdo_modnby1(const uint64_t *limbs, uint64_t nlen, uint64_t d)
{
    if (nlen == 0) return 0;
    int s = __builtin_clzll(d);
    uint64_t nd = d << s;
    uint64_t rnd = ctdo_reciprocal(nd);
    uint64_t r = 0;
    for (uint64_t j = nlen; j-- > 0;) r = ctdo_mod2by1(r, limbs[j], nd, rnd);
    return r % d;
}

static unsigned bit_length(uint64_t n) { return n ? 64u - (unsigned)__builtin_clzll(n) : 0u; }

/* digits of n in base b (modular_arithmetic_detail.hpp:306-326): bases
 * 2,4,8,16,32 count bits; every other base divides (and so gives 1 for n=0). */
uint64_t ctdo_digits_in_base(uint64_t base, uint64_t n)
{
    if (base <= 32 && (base & (base - 1)) == 0) {
        unsigned bpd = bit_length(base) - 1;
        return (bit_length(n) + bpd - 1) / bpd;
    }
    uint64_t res = 1;
    while ((n /= base)) ++res;
    return res;
}

/* ======================================================================== */
/* Prime supply: restates siever.hpp:179-237 with T >= 2 (range [start,end))*/
/* ======================================================================== */

static uint64_t isqrt64(uint64_t n)
{
    uint64_t r = 0;
    for (int b = 31; b >= 0; --b) {
        uint64_t c = r | (1ull << b);
        if (c <= UINT32_MAX && c * c <= n) r = c;
    }
    return r;
}

/* Deterministic Miller-Rabin for n < 2^64 (bases of J. Sinclair). */
static int is_prime_mr(uint64_t n)
{
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (unsigned i = 0; i < sizeof small / sizeof small[0]; ++i) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (unsigned i = 0; i < sizeof bases / sizeof bases[0]; ++i) {
        uint64_t a = bases[i] % n;
        if (a == 0) continue;
        uint64_t x = 1, m = a, e = d;
        while (e) {
            if (e & 1) x = (uint64_t)(((u128)x * m) % n);
            m = (uint64_t)(((u128)m * m) % n);
            e >>= 1;
        }
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; ++r) {
            x = (uint64_t)(((u128)x * x) % n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* Odd primes p <= lim by a plain sieve of Eratosthenes (odd numbers only). */
static uint32_t *small_primes(uint64_t lim, uint64_t *count)
{
    uint64_t half = lim / 2 + 1;
    unsigned char *comp = calloc(half, 1);
    for (uint64_t i = 1; (2 * i + 1) * (2 * i + 1) <= lim; ++i)
        if (!comp[i])
            for (uint64_t j = (2 * i + 1) * (2 * i + 1) / 2; j < half; j += 2 * i + 1) comp[j] = 1;
    uint64_t n = 0;
    for (uint64_t i = 1; i < half && 2 * i + 1 <= lim; ++i) n += !comp[i];
    uint32_t *pr = malloc((n ? n : 1) * sizeof *pr);
    n = 0;
    for (uint64_t i = 1; i < half && 2 * i + 1 <= lim; ++i)
        if (!comp[i]) pr[n++] = (uint32_t)(2 * i + 1);
    free(comp);
    *count = n;
    return pr;
}

typedef void (*prime_sink)(uint64_t p, void *user);

/* Sieving primes for ranges ending at or below `end`. */
typedef struct { uint32_t *pr; uint64_t np; int use_mr; } sieve_base;

static sieve_base sieve_base_make(uint64_t start, uint64_t end)
{
    sieve_base b = {NULL, 0, 0};
    uint64_t lim = end > 1 ? isqrt64(end - 1) : 0;
    if (lim > (1ull << 26) && end - start <= (1ull << 22)) {
        b.use_mr = 1;
        return b;
    }
    b.pr = small_primes(lim, &b.np);
    return b;
}

/* Visits the odd primes of [start, end), start >= 3, in ascending order. */
static void for_odd_primes(uint64_t start, uint64_t end, const sieve_base *sb,
                           prime_sink sink, void *user)
{
    if (start >= end) return;
    if (sb->use_mr) {
        for (uint64_t n = start | 1; n < end && n >= start; n += 2)
            if (is_prime_mr(n)) sink(n, user);
        return;
    }
    const uint64_t SEG = 1ull << 18; /* odd numbers per segment */
    unsigned char *comp = malloc(SEG);
    uint64_t lo = start | 1; /* first odd >= start */
    while (lo < end && lo >= start) {
        uint64_t cnt = (end - lo + 1) / 2;
        if (cnt > SEG) cnt = SEG;
        memset(comp, 0, cnt);
        uint64_t hi = lo + 2 * (cnt - 1); /* last odd candidate */
        for (uint64_t i = 0; i < sb->np; ++i) {
            uint64_t p = sb->pr[i];
            if (p * p > hi) break;
            uint64_t first = p * p;
            if (first < lo) {
                uint64_t r = lo % p;
                first = r ? lo + (p - r) : lo;
                if (!(first & 1)) first += p;
            }
            for (uint64_t m = first; m <= hi; m += 2 * p) comp[(m - lo) / 2] = 1;
        }
        for (uint64_t i = 0; i < cnt; ++i)
            if (!comp[i]) sink(lo + 2 * i, user);
        if (hi + 2 < hi) break;
        lo = hi + 2;
    }
    free(comp);
}

/* Visits the primes of [start, end) in ascending order. */
static void for_primes(uint64_t start, uint64_t end, prime_sink sink, void *user)
{
    if (start >= end) return;
    if (start <= 2 && end > 2) sink(2, user);
    if (start < 3) start = 3;
    if (start >= end) return;
    sieve_base sb = sieve_base_make(start, end);
    for_odd_primes(start, end, &sb, sink, user);
    free(sb.pr);
}

typedef struct { uint64_t *out, cap, n; } collect_t;

static void collect_sink(uint64_t p, void *u)
{
    collect_t *c = u;
    if (c->n < c->cap) c->out[c->n] = p;
    c->n++;
}

uint64_t ctdo_sieve(uint64_t start, uint64_t end, uint64_t *out, uint64_t cap)
{
    collect_t c = {out, cap, 0};
    for_primes(start, end, collect_sink, &c);
    return c.n;
}

static void count_sink(uint64_t p, void *u) { (void)p; ++*(uint64_t *)u; }

uint64_t ctdo_prime_count(uint64_t start, uint64_t end)
{
    uint64_t n = 0;
    for_primes(start, end, count_sink, &n);
    return n;
}

/* ======================================================================== */
/* Sequence families: src/number_sequences/                                  */
/* ======================================================================== */

typedef struct { uint64_t a, b, c; } fstate; /* family state (<= 3 words) */

/* Smarandache start value as limbs (smarandache.cpp:129-158 returns the same
 * integer through its C_k closed form): exact concatenation, grouped so that
 * each big-integer step multiplies by one word. */
uint64_t ctdo_smarandache_value(uint64_t base, uint64_t n, uint64_t *limbs, uint64_t cap)
{
    uint64_t len = 0, alloc = 16;
    uint64_t *v = calloc(alloc, sizeof *v);
    uint64_t i = 1;
    while (i <= n) {
        uint64_t mult = 1, chunk = 0;
        while (i <= n) {
            uint64_t bd = 1;
            uint64_t d = ctdo_digits_in_base(base, i);
            for (uint64_t k = 0; k < d; ++k) bd *= base;
            if (mult > UINT64_MAX / bd) break;
            chunk = chunk * bd + i;
            mult *= bd;
            ++i;
        }
        /* v = v * mult + chunk */
        u128 carry = chunk;
        for (uint64_t k = 0; k < len; ++k) {
            u128 t = (u128)v[k] * mult + carry;
            v[k] = (uint64_t)t;
            carry = t >> 64;
        }
        if (carry) {
            if (len == alloc) {
                alloc *= 2;
                v = realloc(v, alloc * sizeof *v);
            }
            v[len++] = (uint64_t)carry;
        }
    }
    for (uint64_t k = 0; k < len && k < cap; ++k) limbs[k] = v[k];
    free(v);
    return len;
}

/* Exact remainder of a limb vector (HgInt::mod = mpz_fdiv_ui, hgint.hpp:270-274). */
static uint64_t limbs_mod(const uint64_t *v, uint64_t len, uint64_t d)
{
    u128 r = 0;
    for (uint64_t k = len; k-- > 0;) r = ((r << 64) | v[k]) % d;
    return (uint64_t)r;
}

typedef struct {
    uint32_t seq;
    uint64_t param;
    uint64_t n0;
    const uint64_t *sm_limbs; /* Smarandache S(n0) */
    uint64_t sm_len;
} seqdef;

/* value_mod_2 (number_sequence_helpers.hpp:14-24) per family:
 * mersenne.hpp:69-75, riesel_proth.hpp:74-81, riesel_proth_fixed_n.hpp:117-122,
 * smarandache.hpp:83-95, fibonacci.hpp:49-54, lucas.hpp:40-45,
 * tribonacci.hpp:104-108; the test sequence falls back to value_mod(n, 2). */
uint64_t ctdo_value_mod_2(uint32_t seq, uint64_t param, uint64_t n)
{
    switch (seq) {
    case CTDO_MERSENNE: return n > 0;
    case CTDO_PROTH:
    case CTDO_RIESEL: return n > 0 ? 1 : ((param & 1) ^ 1);
    case CTDO_PROTH_FIXED_N:
    case CTDO_RIESEL_FIXED_N: return 1;
    case CTDO_SMARANDACHE:
        if (!(param & 1)) return n & 1;
        return (n & 3) && ((n & 3) != 3);
    case CTDO_FIBONACCI:
    case CTDO_LUCAS: return n % 3 == 0 ? 0 : 1;
    case CTDO_TRIBONACCI: return (n & 3) >> 1;
    case CTDO_PEANO: return n % 2;
    }
    return 0;
}

/* next_value_mod_2 (number_sequence_helpers.hpp:26-37) per family:
 * mersenne.hpp:98-107, riesel_proth.hpp:83-88, riesel_proth_fixed_n.hpp:135-140,
 * smarandache.hpp:97-112, fibonacci.hpp:56-61, lucas.hpp:47-51,
 * tribonacci.hpp:110-115; the test sequence uses next_value_mod(r, n, 2). */
uint64_t ctdo_next_value_mod_2(uint32_t seq, uint64_t param, uint64_t r, uint64_t n)
{
    switch (seq) {
    case CTDO_MERSENNE:
    case CTDO_PROTH:
    case CTDO_RIESEL:
    case CTDO_PROTH_FIXED_N:
    case CTDO_RIESEL_FIXED_N: return 1;
    case CTDO_SMARANDACHE:
        if (!(param & 1)) return (n & 1) ^ 1;
        return r ^ (n & 1) ^ 1;
    case CTDO_FIBONACCI:
    case CTDO_LUCAS: return (n + 1) % 3 == 0 ? 0 : 1;
    case CTDO_TRIBONACCI: return ((n + 1) & 3) >> 1;
    case CTDO_PEANO: return r + 1 == 2 ? 0 : r + 1;
    }
    return 0;
}

/* Fibonacci fast doubling to {F(n-1), F(n)} mod d, n <= 1 special-cased
 * exactly as fibonacci.hpp:64-88 does (n = 0 yields {0, 0}). */
static fstate fib_state(uint64_t n, const bctx *c)
{
    fstate st = {0, 0, 0};
    uint64_t d = c->d;
    if (n <= 1) {
        st.b = n;
        return st;
    }
    n -= 1;
    uint64_t f0 = 0, f1 = 1;
    for (uint64_t m = 1ull << (bit_length(n) - 1); m; m >>= 1) {
        if (m & n) {
            uint64_t t0 = mulmod_b(submod(doublemod(f1, d), f0, d), f0, c);
            f0 = addmod(mulmod_b(f1, f1, c), mulmod_b(f0, f0, c), d);
            f1 = addmod(f0, t0, d);
        } else {
            uint64_t t0 = f0;
            f0 = mulmod_b(submod(doublemod(f1, d), f0, d), f0, c);
            f1 = addmod(mulmod_b(f1, f1, c), mulmod_b(t0, t0, c), d);
        }
    }
    st.a = f0; /* previous */
    st.b = f1; /* current */
    return st;
}

/* (T(n), T(n+1), T(n+2)) mod d by the doubling of tribonacci.hpp:64-102 */
static fstate trib_state(uint64_t n, const bctx *c)
{
    uint64_t d = c->d;
    uint64_t vn = 0, vn1 = 0, vn2 = 1;
    if (n) {
        for (uint64_t m = 1ull << (bit_length(n) - 1); m; m >>= 1) {
            uint64_t sqr_vn = mulmod_b(vn, vn, c);
            uint64_t sqr_vn1 = mulmod_b(vn1, vn1, c);
            uint64_t two_vn1_vn = mulmod_b(doublemod(vn1, d), vn, c);
            uint64_t tvn = submod(submod(addmod(mulmod_b(doublemod(vn2, d), vn, c), sqr_vn1, d),
                                         two_vn1_vn, d), sqr_vn, d);
            uint64_t tvn1 = submod(addmod(sqr_vn, mulmod_b(doublemod(vn2, d), vn1, c), d), sqr_vn1, d);
            vn2 = addmod(addmod(two_vn1_vn, mulmod_b(vn2, vn2, c), d), sqr_vn1, d);
            vn = tvn;
            vn1 = tvn1;
            if (m & n) {
                uint64_t vn3 = addmod(addmod(vn, vn1, d), vn2, d);
                vn = vn1;
                vn1 = vn2;
                vn2 = vn3;
            }
        }
    }
    fstate st = {vn, vn1, vn2};
    return st;
}

/* S(n0) mod d and the family state, as trial_factoring_impl.hpp:201-228 does. */
static uint64_t seq_init(const seqdef *s, const bctx *c, fstate *st)
{
    uint64_t d = c->d, n = s->n0, r;
    memset(st, 0, sizeof *st);
    switch (s->seq) {
    case CTDO_MERSENNE: /* mersenne.hpp:77-96 */
        if (n <= 1) return (n == 0 || d == 1) ? 0 : 1;
        r = modpow_b(2, n, c);
        return r ? r - 1 : d - 1;
    case CTDO_PROTH: /* riesel_proth.hpp:90-101 */
        r = mulmod_b(modpow_b(2, n, c), s->param, c);
        return r == d - 1 ? 0 : r + 1;
    case CTDO_RIESEL:
        r = mulmod_b(modpow_b(2, n, c), s->param, c);
        return r ? r - 1 : d - 1;
    case CTDO_PROTH_FIXED_N: /* riesel_proth_fixed_n.hpp:88-94,124-133 */
    case CTDO_RIESEL_FIXED_N:
        st->a = modpow_b(2, s->param, c);
        r = mulmod_b(ctdo_mod(n, d, c->mu64), st->a, c);
        if (s->seq == CTDO_PROTH_FIXED_N) return r == d - 1 ? 0 : r + 1;
        return r ? r - 1 : d - 1;
    case CTDO_SMARANDACHE: /* initialize_from_value: HgInt::mod (trial_factoring_impl.hpp:217-218) */
        return limbs_mod(s->sm_limbs, s->sm_len, d);
    case CTDO_FIBONACCI: /* fibonacci.hpp:64-96 */
        *st = fib_state(n, c);
        return st->b;
    case CTDO_LUCAS: { /* lucas.hpp:54-66 */
        fstate f = fib_state(n, c);
        st->a = submod(doublemod(f.b, d), f.a, d);
        st->b = addmod(f.b, doublemod(f.a, d), d);
        return st->b;
    }
    case CTDO_TRIBONACCI: /* tribonacci.hpp:117-131 */
        *st = trib_state(n, c);
        return st->a;
    case CTDO_PEANO: /* tests/peano.hpp:41-82 */
        return n % d;
    }
    return 0;
}

/* Smarandache multiplier base^digits(n+1) (smarandache.hpp:160-198) */
static uint64_t sm_multiplier(uint64_t base, uint64_t n)
{
    uint64_t m = n + 1, result = base;
    while ((m /= base)) result *= base;
    return result;
}

/* S(n+1) mod d from r = S(n) mod d, per family. */
static uint64_t seq_next(const seqdef *s, uint64_t r, uint64_t n, const bctx *c, fstate *st)
{
    uint64_t d = c->d;
    switch (s->seq) {
    case CTDO_MERSENNE: /* mersenne.hpp:110-124 */
        if (r < d - r) {
            uint64_t t = (r << 1) + 1;
            return t == d ? 0 : t;
        }
        return r - (d - r) + 1;
    case CTDO_PROTH: /* riesel_proth.hpp:103-115 */
        r = (r < d - r) ? r << 1 : r - (d - r);
        return r ? r - 1 : d - 1;
    case CTDO_RIESEL:
        r = (r < d - r) ? r << 1 : r - (d - r);
        return r == d - 1 ? 0 : r + 1;
    case CTDO_PROTH_FIXED_N: /* riesel_proth_fixed_n.hpp:142-151 */
    case CTDO_RIESEL_FIXED_N:
        return (r >= d - st->a) ? r - (d - st->a) : r + st->a;
    case CTDO_SMARANDACHE: { /* smarandache.hpp:131-158 */
        uint64_t base = s->param;
        u128 a;
        if ((base & (base - 1)) == 0) {
            unsigned sh = (unsigned)(ctdo_digits_in_base(base, n + 1) * (bit_length(base) - 1));
            a = ((u128)r << sh) + n + 1;
        } else {
            a = (u128)r * sm_multiplier(base, n) + n + 1;
        }
        u128 q = (a >> 64) ? mul128hi_u(a, c->mu128) : ((a * c->mu64) >> 64);
        u128 rr = a - q * d;
        return (uint64_t)(rr >= d ? rr - d : rr);
    }
    case CTDO_FIBONACCI: /* fibonacci.hpp:98-104, lucas.hpp:74-78 */
    case CTDO_LUCAS: {
        uint64_t res = addmod(r, st->a, d);
        st->a = r;
        return res;
    }
    case CTDO_TRIBONACCI: { /* tribonacci.hpp:133-141 */
        uint64_t t3 = addmod(addmod(st->a, st->b, d), st->c, d);
        st->a = st->b;
        st->b = st->c;
        st->c = t3;
        return st->a;
    }
    case CTDO_PEANO: /* tests/peano.hpp:85-117 */
        r = r + 1;
        return r == d ? 0 : r;
    }
    return 0;
}

static int seq_known(uint32_t seq);

int ctdo_residues_last(uint32_t seq, uint64_t param, uint64_t n0, uint64_t count,
                       const uint64_t *d, uint64_t nd, uint64_t *out)
{
    if (!seq_known(seq) || count == 0) return -1;
    seqdef s = {seq, param, n0, NULL, 0};
    uint64_t *limbs = NULL;
    if (seq == CTDO_SMARANDACHE) {
        uint64_t len = ctdo_smarandache_value(param, n0, NULL, 0);
        limbs = malloc((len ? len : 1) * sizeof *limbs);
        ctdo_smarandache_value(param, n0, limbs, len);
        s.sm_limbs = limbs;
        s.sm_len = len;
    }
    for (uint64_t k = 0; k < nd; ++k) {
        if (!(d[k] & 1) || d[k] < 3) {
            free(limbs);
            return -1;
        }
        bctx c = bctx_make(d[k]);
        fstate st;
        uint64_t r = seq_init(&s, &c, &st);
        for (uint64_t i = 1; i < count; ++i) r = seq_next(&s, r, n0 + i - 1, &c, &st);
        out[k] = r;
    }
    free(limbs);
    return 0;
}

uint64_t ctdo_next(uint32_t seq, uint64_t param, uint64_t r, uint64_t n, uint64_t d)
{
    seqdef s = {seq, param, 0, NULL, 0};
    bctx c = bctx_make(d);
    fstate st = {0, 0, 0};
    return seq_next(&s, r, n, &c, &st);
}

int ctdo_state(uint32_t seq, uint64_t param, uint64_t n, uint64_t d, uint64_t out[3])
{
    bctx c = bctx_make(d);
    fstate st;
    switch (seq) {
    case CTDO_FIBONACCI:
        st = fib_state(n, &c);
        out[0] = st.a;
        out[1] = st.b;
        return 2;
    case CTDO_LUCAS: {
        seqdef s = {seq, param, n, NULL, 0};
        (void)seq_init(&s, &c, &st);
        out[0] = st.a;
        out[1] = st.b;
        return 2;
    }
    case CTDO_TRIBONACCI:
        st = trib_state(n, &c);
        out[0] = st.a;
        out[1] = st.b;
        out[2] = st.c;
        return 3;
    case CTDO_PROTH_FIXED_N:
    case CTDO_RIESEL_FIXED_N:
        out[0] = modpow_b(2, param, &c);
        return 1;
    }
    return 0;
}

static int seq_known(uint32_t seq)
{
    return seq <= CTDO_TRIBONACCI || seq == CTDO_PEANO;
}

int ctdo_residues(uint32_t seq, uint64_t param, uint64_t n0, uint64_t count,
                  uint64_t d, uint64_t *out)
{
    if (!seq_known(seq) || !(d & 1) || d < 3) return -1;
    seqdef s = {seq, param, n0, NULL, 0};
    uint64_t *limbs = NULL;
    if (seq == CTDO_SMARANDACHE) {
        uint64_t len = ctdo_smarandache_value(param, n0, NULL, 0);
        limbs = malloc((len ? len : 1) * sizeof *limbs);
        ctdo_smarandache_value(param, n0, limbs, len);
        s.sm_limbs = limbs;
        s.sm_len = len;
    }
    bctx c = bctx_make(d);
    fstate st;
    if (count) {
        uint64_t r = seq_init(&s, &c, &st);
        out[0] = r;
        for (uint64_t i = 1; i < count; ++i) {
            r = seq_next(&s, r, n0 + i - 1, &c, &st);
            out[i] = r;
        }
    }
    free(limbs);
    return 0;
}

/* ======================================================================== */
/* The engine: src/trial_factoring_impl.hpp:193-366                         */
/* ======================================================================== */

typedef struct {
    const ctdo_job *job;
    seqdef s;
    uint32_t *counts;
    uint64_t *factors;
} tfctx;

/* factors_buffer::push_factor (factors_buffer.hpp:110-121): atomic count,
 * store while below the per-index maximum. */
static void push_factor(const tfctx *t, uint64_t idx, uint64_t p)
{
    uint32_t pos = __atomic_fetch_add(&t->counts[idx], 1u, __ATOMIC_RELAXED);
    if (pos < t->job->max_factors) t->factors[idx * t->job->max_factors + pos] = p;
}

/* trial_factorer::process_chunk body for one prime (trial_factoring_impl.hpp:199-246) */
static void tf_prime(uint64_t p, void *u)
{
    const tfctx *t = u;
    const ctdo_job *j = t->job;
    bctx c = bctx_make(p);
    fstate st;
    uint64_t r = seq_init(&t->s, &c, &st);
    if (!r) push_factor(t, 0, p);
    for (uint64_t n = j->n0, ne = j->n1 - 1; n < ne; ++n) {
        r = seq_next(&t->s, r, n, &c, &st);
        if (!r) push_factor(t, n + 1 - j->n0, p);
    }
}

typedef struct {
    tfctx *t;
    uint64_t lo, hi, seg;
    uint64_t next; /* shared segment cursor */
    sieve_base sb;
} pool_t;

static void *tf_worker(void *u)
{
    pool_t *pl = u;
    for (;;) {
        uint64_t k = __atomic_fetch_add(&pl->next, 1, __ATOMIC_RELAXED);
        uint64_t a = pl->lo + k * pl->seg;
        if (a >= pl->hi || a < pl->lo) break;
        uint64_t b = (pl->hi - a > pl->seg) ? a + pl->seg : pl->hi;
        for_odd_primes(a, b, &pl->sb, tf_prime, pl->t);
    }
    return NULL;
}

int ctdo_trial_factor(const ctdo_job *job, uint32_t threads, uint32_t *counts, uint64_t *factors)
{
    if (!seq_known(job->seq)) return -1;
    if (job->n1 <= job->n0 || job->f1 <= job->f0) return 0; /* trial_factoring_impl.hpp:380-387 */
    tfctx t = {job, {job->seq, job->param, job->n0, NULL, 0}, counts, factors};
    uint64_t *limbs = NULL;
    if (job->seq == CTDO_SMARANDACHE) { /* first_value (trial_factoring_impl.hpp:179-191) */
        uint64_t len = ctdo_smarandache_value(job->param, job->n0, NULL, 0);
        limbs = malloc((len ? len : 1) * sizeof *limbs);
        ctdo_smarandache_value(job->param, job->n0, limbs, len);
        t.s.sm_limbs = limbs;
        t.s.sm_len = len;
    }
    uint64_t f0 = job->f0;
    /* prime 2 through the mod-2 helpers (trial_factoring_impl.hpp:312-325) */
    if (f0 <= 2 && job->f1 > 2) {
        f0 = 3;
        uint64_t r = ctdo_value_mod_2(job->seq, job->param, job->n0);
        if (!r) push_factor(&t, 0, 2);
        for (uint64_t n = job->n0, ne = job->n1 - 1; n < ne; ++n) {
            r = ctdo_next_value_mod_2(job->seq, job->param, r, n);
            if (!r) push_factor(&t, n + 1 - job->n0, 2);
        }
    }
    if (f0 < 3) f0 = 3; /* no odd prime below 3: [0, 2) and [1, 2) hold no prime at all */
    if (f0 < job->f1) {
        if (threads < 1) threads = 1;
        pool_t pl = {&t, f0, job->f1, 1ull << 22, 0, sieve_base_make(f0, job->f1)};
        if (threads == 1) {
            tf_worker(&pl);
        } else {
            pthread_t *th = malloc(threads * sizeof *th);
            for (uint32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, tf_worker, &pl);
            for (uint32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
            free(th);
        }
        free(pl.sb.pr);
    }
    free(limbs);
    return 0;
}

uint64_t ctdo_job_primes(const ctdo_job *job)
{
    if (job->f1 <= job->f0) return 0;
    return ctdo_prime_count(job->f0, job->f1);
}
