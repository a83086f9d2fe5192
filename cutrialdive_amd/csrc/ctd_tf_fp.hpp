// ctd_tf_fp.hpp -- residue propagation in exact FP64 arithmetic, p < 2^48.
//
// On gfx950 a 64-bit integer add or subtract is two VALU instructions (low
// word, then high word with carry) while an FP64 add or fused multiply-add is
// one, at the same full rate (the engine's probe: v_add_f64 37.8 T, v_fma_f64
// 34.2 T lane-ops/s).  Below 2^53 FP64 arithmetic on integers is exact, so a
// residue can live in a double:
//
//   * values are integers kept in the signed range |x| <= 0.75 p;
//   * red(v) = v - p round(v / p), with round(v/p) = fma(v, 1/p, 1.5 2^52) - 1.5 2^52
//     (exact integer result; 1/p is only an estimate, so round() may be off
//     by one near a half-integer, which leaves |red(v)| <= 0.75 p);
//   * mul(a, b) = hi - p round(hi/p) + lo with hi = ab rounded and
//     lo = fma(a, b, -hi) its exact error (Dekker's product): the value is
//     congruent to ab, and every operation is exact because each result fits
//     in 53 bits (|ab| < 2^100, |ab/p| < 2^50, |hi - p round(hi/p)| < 2p).
//
// Every residue class has one representative with |x| < p, so after a
// reduction the zero test is x == 0 and the doubling families test x == +-1.
//
// Stride forms (doubling and fixed-n families).  When the reduction is exact
// (round(v/p) correct, which holds when |v/p| 2^-52 < 1/(2p), i.e. |v| < 2^51:
// the distance of v/p from a half-integer is at least 1/(2p) for odd p) every
// class has the single representative |x| <= (p-1)/2.  Then the S indices
// after the state x_n are tested by S compares of x_n against precomputed
// targets, and the state advances S indices at once:
//   doubling  x_{n+j} = x_n 2^j:   hit at n+j  <=>  x_n == t 2^-j,
//             S = 128 (then 64 and 16 for the tail), x <- x (2^S mod p), an
//             exact product (mulc_canon, 9 instructions);
//   fixed n   x_{n+j} = x_n + j c: hit at n+j  <=>  x_n == -j c, S = 64,
//             x <- red(x + S c)  (4 instructions).
// The compares use the low 32-bit words of x + M and of the targets
// (low_word), two target words per register pair, and a block whose words
// match is replayed one index at a time with the exact value.
//
// Per index this costs (VALU instructions, compare included):
//
//   doubling x = 2x           1.08 (integer W64 form: 7)
//   fixed n  x = x + c        1.08 (6) when stepped; by default the closed
//                             form of ctd_tf_fast.hpp solves once per prime
//   fibonacci / lucas         1.30 / 1.27 theta stride form, 5 pair form (6)
//   tribonacci                3 + 7/L: 3.64 (p < 2^43, L = 11) to 4.4 (p < 2^48, L = 5); 11.25 above
//   smarandache r m + n + 1   7   (12.25), for p > 8 n1; 5 when p (0.75 m + 1) < 2^53
//                             (short step) or (p / 2 + n1) o < 2^53 (shifted step,
//                             o the odd part of m)
//
// and the per-prime set-up (powers, doubling formulas, Smarandache jump) runs
// on 6-instruction multiplies instead of ~40-instruction Montgomery products.
#pragma once

// included by ctd_tf_fast.hpp after record_hits and CTD_UNROLLED_LOOP

namespace ctd {

constexpr uint64_t FP_PMAX = 1ull << 48;   // every prime of the wave must be below
// the Smarandache short form needs p (0.75 m + 1) below 2^53 (margin for the rounded product)
constexpr double SM_SHORT_LIMIT = 9007199254740992.0 - 1073741824.0;
// above 1/2 + 2^-30: the reduced part of a Smarandache state is at most this times p
constexpr double SM_EXACT_HALF = 0.5000001;

struct Fmod {
    double p, invp;
    static constexpr double MAGIC = 6755399441055744.0; // 1.5 * 2^52: adding it rounds to an integer

    __device__ __forceinline__ void init(uint64_t pi)
    {
        p = (double)pi;
        invp = 1.0 / p;
    }
    // round(v / p) for |v / p| < 2^50, possibly off by one near a half-integer
    __device__ __forceinline__ double quot(double v) const { return fma(v, invp, MAGIC) - MAGIC; }
    __device__ __forceinline__ double red(double v) const { return fma(-quot(v), p, v); }
    __device__ __forceinline__ double mul(double a, double b) const
    {
        const double hi = a * b;
        const double lo = fma(a, b, -hi);
        return fma(-quot(hi), p, hi) + lo;
    }
    __device__ __forceinline__ double add(double a, double b) const { return red(a + b); }
    __device__ __forceinline__ double sub(double a, double b) const { return red(a - b); }
    // the residue of an integer below 2^53
    __device__ __forceinline__ double of(uint64_t x) const { return red((double)x); }
    // 2^e (left-to-right square and double)
    __device__ double pow2(uint64_t e) const
    {
        double x = 1.0;
        if (e == 0) return x;
        for (int b = 63 - __clzll(e); b >= 0; --b) {
            x = mul(x, x);
            if ((e >> b) & 1) x = red(x + x);
        }
        return x;
    }
    __device__ double pow(double x, uint64_t e) const
    {
        double r = 1.0;
        while (e) {
            if (e & 1) r = mul(r, x);
            x = mul(x, x);
            e >>= 1;
        }
        return r;
    }
    // canonical residue in [0, p) of |x| < 2^52
    __device__ __forceinline__ uint64_t canon(double x, uint64_t pi) const
    {
        const int64_t t = (int64_t)red(x);
        return t < 0 ? (uint64_t)(t + (int64_t)pi) : (uint64_t)t;
    }
};

// (F(m), F(m+1)) by fast doubling
__device__ inline void fib_pair_fp(const Fmod &F, uint64_t m, double &fa, double &fb)
{
    double a = 0.0, b = 1.0;
    if (m) {
        for (int bit = 63 - __clzll(m); bit >= 0; --bit) {
            const double c = F.mul(a, F.sub(b + b, a));        // F(2k)
            const double d = F.add(F.mul(a, a), F.mul(b, b));   // F(2k+1)
            if ((m >> bit) & 1) {
                a = d;
                b = F.add(c, d);
            } else {
                a = c;
                b = d;
            }
        }
    }
    fa = a;
    fb = b;
}

// (T(n), T(n+1), T(n+2)), the doubling identities of fib_pair's tribonacci
// counterpart trib_triple (ctd_families.hpp)
__device__ inline void trib_triple_fp(const Fmod &F, uint64_t n, double &ta, double &tb, double &tc)
{
    double a = 0.0, b = 0.0, c = 1.0;
    if (n) {
        for (int bit = 63 - __clzll(n); bit >= 0; --bit) {
            const double aa = F.mul(a, a), bb = F.mul(b, b), cc = F.mul(c, c);
            const double ab2 = F.mul(a, b + b), ac2 = F.mul(a, c + c), bc2 = F.mul(b, c + c);
            const double t0 = F.red(ac2 + bb - ab2 - aa); // T(2k)
            const double t1 = F.red(aa + bc2 - bb);       // T(2k+1)
            const double t2 = F.red(ab2 + cc + bb);       // T(2k+2)
            if ((n >> bit) & 1) {
                a = t1;
                b = t2;
                c = F.red(t0 + t1 + t2);
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

// Sm_b(n) mod p by affine block jumps (the same decomposition as sm_value in
// ctd_families.hpp): appending t consecutive d-digit numbers m0..m0+t-1 maps S
// to B^t S + m0 U_t + V_t.  Below SM_DIRECT_MAX the n steps themselves are
// cheaper (about 10 instructions each against about 60 per bit of t).
constexpr uint64_t SM_DIRECT_MAX = 32;

__device__ inline double sm_value_fp(const Fmod &F, uint64_t b, uint64_t n)
{
    if (n == 0) return 0.0;
    if (n < SM_DIRECT_MAX) {
        double S = 0.0, BF = F.of(b);
        uint64_t nxt = b; // the first number with one more digit
        for (uint64_t i = 1; i <= n; ++i) {
            if (i == nxt) {
                BF = F.mul(BF, (double)b);
                nxt *= b;
            }
            S = F.red(F.mul(S, BF) + (double)i);
        }
        return S;
    }
    const double bd = (double)b;
    double S = 0.0;
    uint64_t m0 = 1;
    double m0F = 1.0, BF = F.of(b);
    for (;;) {
        const uint64_t m1 = sat_mul(m0, b);
        const uint64_t last = n < m1 - 1 ? n : m1 - 1;
        const uint64_t t = last - m0 + 1;
        double P = 1.0, U = 0.0, V = 0.0, tF = 0.0;
        for (int bit = 63 - __clzll(t); bit >= 0; --bit) {
            V = F.red(F.mul(P, V) + F.mul(tF, U) + V);
            U = F.mul(U, P + 1.0);
            P = F.mul(P, P);
            tF = F.red(tF + tF);
            if ((t >> bit) & 1) {
                V = F.add(F.mul(BF, V), tF);
                U = F.add(F.mul(U, BF), 1.0);
                P = F.mul(P, BF);
                tF = F.add(tF, 1.0);
            }
        }
        S = F.red(F.mul(P, S) + F.mul(m0F, U) + V);
        if (last == n || m1 == UINT64_MAX) break;
        m0 = m1;
        m0F = F.mul(m0F, bd);
        BF = F.mul(BF, bd);
    }
    return S;
}

// Z_p[phi] / (phi^2 - phi - 1), elements u + v phi with reduced coordinates.
// theta = phi / psi = -(1 + phi) (psi = 1 - phi, phi psi = -1), so that
// F(n) (phi - psi) = psi^n (theta^n - 1) and L(n) = psi^n (theta^n + 1).  For
// p != 2, 5 both phi - psi (its square is 5) and psi are units, hence
//   F(n) = 0 mod p  <=>  theta^n = 1,   L(n) = 0 mod p  <=>  theta^n = -1.
struct Rphi {
    double u, v;
};

// x theta = -(u + v) - (u + 2v) phi
__device__ __forceinline__ Rphi rphi_mul_theta(const Fmod &F, Rphi x)
{
    const double s = x.u + x.v;
    return {F.red(-s), F.red(-s - x.v)};
}
// x y = (xu yu + xv yv) + (xu yv + xv yu + xv yv) phi
__device__ __forceinline__ Rphi rphi_mul(const Fmod &F, Rphi x, Rphi y)
{
    const double a = F.mul(x.u, y.u), b = F.mul(x.v, y.v);
    const double c = F.mul(F.red(x.u + x.v), F.red(y.u + y.v)); // (xu + xv)(yu + yv)
    return {F.red(a + b), F.red(c - a)};
}
// x times the constant g = (gu, gv), with gw = gu + gv reduced
__device__ __forceinline__ Rphi rphi_mul_const(const Fmod &F, Rphi x, double gu, double gv, double gw)
{
    return {F.red(F.mul(x.u, gu) + F.mul(x.v, gv)), F.red(F.mul(x.u, gv) + F.mul(x.v, gw))};
}
// theta^e
__device__ inline Rphi rphi_theta_pow(const Fmod &F, uint64_t e)
{
    Rphi x = {1.0, 0.0};
    if (e == 0) return x;
    for (int b = 63 - __clzll(e); b >= 0; --b) {
        x = rphi_mul(F, x, x);
        if ((e >> b) & 1) x = rphi_mul_theta(F, x);
    }
    return x;
}

// Indices per reduction of the theta stride form: THETA_STRIDE compares of
// 32-bit target words against 33 VALU instructions of the ring product, so
// 1 + 33 / THETA_STRIDE per index.  Each target word takes a register: with
// 128 words the Lucas kernel needs 166 registers (3 waves per SIMD) and the
// Fibonacci kernel 172 (2 waves, which it runs 9 % slower than 3); 112 keeps
// Fibonacci at 156 (3 waves).  Measured on the mixed batch's jobs against 64
// words (106 registers, 4 waves): Lucas 14 % faster with 128, Fibonacci 10 %
// with 112.  THETA_MIN_STEPS is the shortest work item that takes the form
// (its set-up is about 60 ring products and the targets).
template <uint32_t FAM>
constexpr int THETA_STRIDE = FAM == FAM_LUCAS ? 128 : 112;
constexpr uint64_t THETA_MIN_STEPS = 1024;
// target words of the doubling families' word stride
constexpr int DOUBLING_WORDS = 128;

// One block of the Tribonacci lazy form: L indices from the reduced state
// (a, b, c) = (T(n), T(n+1), T(n+2)), |.| <= 0.75 p.  The terms
// x[j] = T(n+3+j) are formed unreduced: T(n+3) = a + b + c, then
// T(k+4) = 2 T(k+3) - T(k) (subtract T(k+3) = T(k+2) + T(k+1) + T(k) from
// T(k+4) = T(k+3) + T(k+2) + T(k+1)), one FMA per term, exact while the term
// fits in 53 bits.  In terms of the state, T(n+3+j) = A a + B b + C c with
// A + B + C = 3, 5, 9, 17, 31, 57, 105, 193, 355, 653, 1201 for j = 0..10, so
// the last term of a block of L stays below 2^53 with L = 11 for p < 2^43
// (|x[10]| <= 901 p), L = 10 for p < 2^44 (490 p), L = 8 for p < 2^45
// (145 p), L = 7 for p < 2^46 (79 p), L = 6 for p < 2^47 (43 p) and L = 5
// for p < 2^48 (24 p).
//
// The terms that do not become state, x[0..L-4], are tested by a filter of
// one FMA and one 32-bit compare: y = x / p + 1.5 2^20 rounded lies in
// (2^20, 2^21) (|x / p| <= 901), where a double's low word holds the
// fraction of y to 2^-32.  For x = i p, x (1/p rounded) = i (1 + d) with
// |d| <= 2^-53, within 901 2^-53 < 2^-33 of the integer i, so y rounds to
// i + 1.5 2^20 and its low word is 0: the filter never misses a hit.  It
// also passes the rare x within about p 2^-33 of a multiple of p (a 2^-32
// chance for a term that is not a multiple; none for p < 2^32), so a block
// whose filter masks are not empty tests those terms again exactly.
// The last three terms are reduced for the next state: with q the biased
// rounded quotient, w = fma(q, -p, M p) = -round(x/p) p exactly
// (|w| < 2^53), x = 0 mod p <=> x == -w, and x + w is the reduced value
// (|.| <= p/2 + tiny).  T(n+1) and T(n+2) are reduced already and compare
// with 0.  3 + 7/L instructions per index.
constexpr double TRIB_FILTER_BIAS = 1572864.0; // 1.5 * 2^20

__device__ __forceinline__ double trib_neg_round(const Fmod &F, double pM, double x)
{
    return fma(fma(x, F.invp, Fmod::MAGIC), -F.p, pM);
}

template <int L>
__device__ __forceinline__ void trib_lazy_block(const Fmod &F, double pM, double &a, double &b, double &c,
                                                double (&x)[L], uint64_t (&mk)[L])
{
    x[0] = a + (b + c);
    x[1] = fma(2.0, x[0], -a);
    x[2] = fma(2.0, x[1], -b);
    x[3] = fma(2.0, x[2], -c);
    _Pragma("unroll")
    for (int j = 4; j < L; ++j) x[j] = fma(2.0, x[j - 1], -x[j - 4]);
    mk[0] = __ballot(b == 0.0);
    mk[1] = __ballot(c == 0.0);
    _Pragma("unroll")
    for (int j = 0; j < L - 3; ++j) {
        const double y = fma(x[j], F.invp, TRIB_FILTER_BIAS);
        mk[2 + j] = __ballot((uint32_t)__double_as_longlong(y) == 0u);
    }
    const double w0 = trib_neg_round(F, pM, x[L - 3]);
    const double w1 = trib_neg_round(F, pM, x[L - 2]);
    const double w2 = trib_neg_round(F, pM, x[L - 1]);
    mk[L - 1] = __ballot(x[L - 3] == -w0);
    a = x[L - 3] + w0;
    b = x[L - 2] + w1;
    c = x[L - 1] + w2;
}

// The Tribonacci lazy-form loop over `steps` indices after idx (the index of
// the state a = T(n)); blocks of L, then single reduced steps.
template <int L>
__device__ __forceinline__ void trib_lazy_loop(const JobArgs &J, const Fmod &F, uint64_t p, double &a, double &b,
                                               double &c, uint64_t &idx, uint64_t steps)
{
    const double pM = F.p * Fmod::MAGIC;
    const uint32_t steps32 = (uint32_t)steps;
    uint32_t k = 0;
    for (; k + L <= steps32; k += L) {
        double x[L];
        uint64_t mk[L];
        trib_lazy_block<L>(F, pM, a, b, c, x, mk);
        uint64_t any = 0;
        _Pragma("unroll")
        for (int j = 0; j < L; ++j) any |= mk[j];
        if (any) {
            // the filtered terms again, exactly
            _Pragma("unroll")
            for (int j = 0; j < L - 3; ++j)
                if (mk[2 + j]) mk[2 + j] = __ballot(x[j] == -trib_neg_round(F, pM, x[j]));
            const uint32_t lane = __lane_id();
            const uint64_t at = hit_base(idx, k);
            for (int j = 0; j < L; ++j)
                if (mk[j]) record_hits_mask(J, at + 1 + j, p, (mk[j] >> lane) & 1, mk[j]);
        }
    }
    idx += k;
    for (; k < steps32; ++k) {
        const double t = F.red(a + b + c);
        a = b;
        b = c;
        c = t;
        ++idx;
        record_hits(J, idx, p, a == 0.0);
    }
}

// v in a vector register, opaque to the compiler: a rare replay path that
// starts from it cannot reuse the values (wave masks) its loop computed from
// v, which would otherwise stay live across the loop.
__device__ __forceinline__ double opaque_vgpr(double v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// x C mod p, exact and canonical (|result| <= (p-1)/2), for |x|, |C| <= p/2:
// hi + lo = x C exactly (Dekker), hi - round(hi/p) p is exact and small, and
// the final reduction of a value below 2^51 rounds its quotient correctly.
__device__ __forceinline__ double mulc_canon(const Fmod &F, double x, double C)
{
    const double hi = x * C;
    const double lo = fma(x, C, -hi);
    return F.red(fma(-F.quot(hi), F.p, hi) + lo);
}

// t / 2 mod p, canonical, for canonical t: t / 2 when t is even, else
// (t - sign(t) p) / 2 (exact: f is 0 or 1/2).
__device__ __forceinline__ double half_canon(const Fmod &F, double t)
{
    const double h = 0.5 * t;
    const double f = h - floor(h);
    return fma(-f, copysign(F.p, t), h);
}

// The low 32-bit word of v + M (M the rounding constant): for an integer
// |v| < 2^51 the sum is exact and its low word is v's low 32 bits, so equal
// values give equal words.
__device__ __forceinline__ uint32_t low_word(double v)
{
    return (uint32_t)__double_as_longlong(v + Fmod::MAGIC);
}

// Compares against 16 targets that cost the scalar unit 3 instructions instead
// of 16: each v_cmpx_ne removes the lanes whose state equals its target from
// EXEC, and the lanes that left (saved EXEC xor EXEC) are the wave's matches;
// EXEC is restored inside the same statement.  (A compare per target into a
// lane mask costs the scalar unit an OR per compare, and the scalar unit
// issues for all four SIMDs of a CU.)  On gfx950 the 32-bit form writes VCC
// as well, and a VALU instruction reads the EXEC a v_cmpx wrote without wait
// states.
__device__ __forceinline__ uint64_t words_match16(uint32_t w, const uint64_t *Tw)
{
    uint64_t save, gone;
    asm volatile("s_mov_b64 %[s], exec\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[t0]\n\tv_cmpx_ne_u32 vcc, %[w], %[t1]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[t2]\n\tv_cmpx_ne_u32 vcc, %[w], %[t3]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[t4]\n\tv_cmpx_ne_u32 vcc, %[w], %[t5]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[t6]\n\tv_cmpx_ne_u32 vcc, %[w], %[t7]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[t8]\n\tv_cmpx_ne_u32 vcc, %[w], %[t9]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[ta]\n\tv_cmpx_ne_u32 vcc, %[w], %[tb]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[tc]\n\tv_cmpx_ne_u32 vcc, %[w], %[td]\n\t"
                 "v_cmpx_ne_u32 vcc, %[w], %[te]\n\tv_cmpx_ne_u32 vcc, %[w], %[tf]\n\t"
                 "s_xor_b64 %[g], %[s], exec\n\t"
                 "s_mov_b64 exec, %[s]"
                 : [s] "=&s"(save), [g] "=&s"(gone)
                 : [w] "v"(w), [t0] "v"((uint32_t)Tw[0]), [t1] "v"((uint32_t)(Tw[0] >> 32)),
                   [t2] "v"((uint32_t)Tw[1]), [t3] "v"((uint32_t)(Tw[1] >> 32)), [t4] "v"((uint32_t)Tw[2]),
                   [t5] "v"((uint32_t)(Tw[2] >> 32)), [t6] "v"((uint32_t)Tw[3]), [t7] "v"((uint32_t)(Tw[3] >> 32)),
                   [t8] "v"((uint32_t)Tw[4]), [t9] "v"((uint32_t)(Tw[4] >> 32)), [ta] "v"((uint32_t)Tw[5]),
                   [tb] "v"((uint32_t)(Tw[5] >> 32)), [tc] "v"((uint32_t)Tw[6]), [td] "v"((uint32_t)(Tw[6] >> 32)),
                   [te] "v"((uint32_t)Tw[7]), [tf] "v"((uint32_t)(Tw[7] >> 32))
                 : "vcc", "scc");
    return gone;
}

// nb blocks of B doubling indices.  Tw packs the words of the targets
// t 2^-(j+1) (j < 2 N), two per register pair; x advances B indices by one
// exact product with C = 2^B mod p.  A block costs the word of x, B compares
// and 9 instructions of product; a block with a word match is replayed one
// index at a time, exactly.
template <int B, int N>
__device__ __forceinline__ double doubling_words(const JobArgs &J, uint64_t p, uint64_t &idx, uint32_t nb,
                                                 double x, const uint64_t (&Tw)[N], const Fmod &F, double C,
                                                 double tgt)
{
    static_assert(B <= 2 * N, "a block needs one target per index");
    for (uint32_t i = 0; i < nb; ++i) {
        const uint32_t xw = low_word(x);
        uint64_t any = 0;
        _Pragma("unroll")
        for (int q = 0; q < B / 16; ++q) any |= words_match16(xw, Tw + 8 * q);
        if (any) {
            const uint32_t lane = __lane_id();
            const uint64_t at = hit_base(idx, i * (uint32_t)B);
            double y = opaque_vgpr(x); // no reuse of the loop's masks (they would stay live)
            for (int j = 0; j < B; ++j) {
                y = F.red(y + y);
                const uint64_t m = __ballot(y == tgt);
                if (m) record_hits_mask(J, at + 1 + (uint32_t)j, p, (m >> lane) & 1, m);
            }
        }
        x = mulc_canon(F, x, C);
    }
    idx += (uint64_t)nb * B;
    return x;
}

// nb blocks of B fixed-n indices (x advances c per index, cB = B c per
// block; Tw packs the words of the targets -(j+1) c).  A canonical x equal to
// a target has the target's word; a word match is replayed one index at a
// time, exactly.
template <int B>
__device__ __forceinline__ double fixedn_blocks(const JobArgs &J, uint64_t p, uint64_t &idx, uint32_t nb, double x,
                                                const uint64_t (&Tw)[32], const Fmod &F, double c, double cB)
{
    for (uint32_t i = 0; i < nb; ++i) {
        const uint32_t xw = low_word(x);
        uint64_t any = 0;
        _Pragma("unroll")
        for (int q = 0; q < B / 16; ++q) any |= words_match16(xw, Tw + 8 * q);
        if (any) {
            const uint32_t lane = __lane_id();
            const uint64_t at = hit_base(idx, i * (uint32_t)B);
            double y = opaque_vgpr(x);
            for (int j = 0; j < B; ++j) {
                y = F.red(y + c);
                const uint64_t m = __ballot(y == 0.0);
                if (m) record_hits_mask(J, at + 1 + (uint32_t)j, p, (m >> lane) & 1, m);
            }
        }
        x = F.red(x + cB);
    }
    idx += (uint64_t)nb * B;
    return x;
}

// A wave-uniform value moved to scalar registers (the compiler keeps values
// derived from FP64 arithmetic in vector registers even when they are uniform).
__device__ __forceinline__ uint64_t scalar_u64(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The largest i with i o < 2^(e+1), for an odd o < 2^21 and e <= 52, given
// inv_o = 1/o rounded.  X = 2^(e+1) - 1 is exact in a double and X inv_o is
// within 2^-52 X/o < 1 of X/o (o >= 3; exact for o = 1), so the truncated
// quotient is off by at most one and one exact check corrects it.
__device__ __forceinline__ uint64_t binade_last(uint32_t e, uint64_t o, double inv_o)
{
    const uint64_t X = (2ull << e) - 1;
    uint64_t q = scalar_u64((uint64_t)((double)X * inv_o));
    if (q * o > X) --q;
    else if ((q + 1) * o <= X) ++q;
    return q;
}

// Smarandache index steps.  step(v, s) advances the lane state v and the
// wave-uniform scalar sequence s (a double's bit pattern, stepped on the
// scalar unit) by one index; the step numbered kk (from 0) tests index
// at0 + kk by v == 0 afterwards.  With PRODUCT the steps first run in blocks
// of 16 with one wave-wide test per block: every state is a nonzero integer of
// size below 2^48 unless it is a hit, so the product of a block's 16 states is
// 0 exactly when one of them is (at most 2^768: no overflow; at least 1 in
// size: no underflow).  The product costs the VALU what the compares cost,
// but the scalar unit sees one mask per block instead of one per index.  A
// block with a zero product is replayed from its first state with one compare
// per index.  (Measured: the 4-instruction shifted step gains from it, the
// longer steps do not, so they take blocks of 4 with per-index masks.)  The
// K primes of a lane share s and the loop control.
template <int K, bool PRODUCT, typename Step>
__device__ __forceinline__ void sm_blocks(const JobArgs &J, const uint64_t (&p)[K], uint64_t at0, uint32_t cnt,
                                          double (&v)[K], uint64_t &s, Step step)
{
    uint32_t kk = 0;
    for (; PRODUCT && kk + 16 <= cnt; kk += 16) {
        double v0[K], prod[K];
        _Pragma("unroll")
        for (int q = 0; q < K; ++q) v0[q] = v[q];
        const uint64_t s0 = s;
        step(v, s);
        _Pragma("unroll")
        for (int q = 0; q < K; ++q) prod[q] = v[q];
        _Pragma("unroll")
        for (int j = 1; j < 16; ++j) {
            step(v, s);
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) prod[q] *= v[q];
        }
        bool zero = false;
        _Pragma("unroll")
        for (int q = 0; q < K; ++q) zero |= prod[q] == 0.0;
        if (__ballot(zero)) {
            const uint64_t at = hit_base(at0, kk);
            double y[K];
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) y[q] = opaque_vgpr(v0[q]);
            uint64_t sy = s0;
            for (int j = 0; j < 16; ++j) {
                step(y, sy);
                _Pragma("unroll")
                for (int q = 0; q < K; ++q) {
                    const uint64_t mk = __ballot(y[q] == 0.0);
                    if (mk) record_hits_mask(J, at + (uint32_t)j, p[q], (mk >> __lane_id()) & 1, mk);
                }
            }
        }
    }
    // the rest: blocks of 4 with one scalar branch, then single steps
    for (; kk + 4 <= cnt; kk += 4) {
        uint64_t mk[4][K];
        uint64_t any = 0;
        _Pragma("unroll")
        for (int j = 0; j < 4; ++j) {
            step(v, s);
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) {
                mk[j][q] = __ballot(v[q] == 0.0);
                any |= mk[j][q];
            }
        }
        if (any) {
            const uint32_t lane = __lane_id();
            const uint64_t at = hit_base(at0, kk);
            _Pragma("unroll")
            for (int j = 0; j < 4; ++j) {
                _Pragma("unroll")
                for (int q = 0; q < K; ++q)
                    if (mk[j][q]) record_hits_mask(J, at + (uint32_t)j, p[q], (mk[j][q] >> lane) & 1, mk[j][q]);
            }
        }
    }
    for (; kk < cnt; ++kk) {
        step(v, s);
        _Pragma("unroll")
        for (int q = 0; q < K; ++q) {
            const uint64_t mk = __ballot(v[q] == 0.0);
            if (mk) record_hits_mask(J, hit_base(at0, kk), p[q], (mk >> __lane_id()) & 1, mk);
        }
    }
}

// The Smarandache FP64 form of Fast<SMARANDACHE>::run for K primes per lane
// (every prime of the wave below FP_PMAX and above 8 n1); fin receives the
// plain residues S(ne-1) mod p for the test hook.  The K primes share the
// segment choices (each form's bound must hold for all of them), the scalar
// sequences and the loop control.
//
    // r = S(n); step r = r m + c with m = b^digits(n+1) < 2^21 (host flag
    // JOB_SM_NARROW) and c = n+1 < 2^32.  Write p = ph + pl with ph the top
    // 32 significant bits of p and let k = round(r m / p), |k| < 2^21:
    //   r' = (r m - k ph) + c - k pl.
    // k ph is exact (at most 21 + 32 significant bits); r m - k ph =
    // (r m - k p) + k pl lies below 2^53, so one fma computes it exactly;
    // the rest is small.  |r m - k p| <= (1/2 + 2^-31) p, so |r'| <= 0.75 p
    // whenever p > 8 n1 (checked per wave), and S(n) = 0 mod p <=> r == 0.
    // c is wave-uniform and lives in scalar registers as the bit pattern
    // of the double: inside one binade [2^e, 2^(e+1)) consecutive integers
    // differ by 2^(52-e) in the pattern (and the carry into the exponent at
    // 2^(e+1) still gives the exact value), so an integer add steps it and
    // the VALU does 7 instructions per index, the compare included.
    //
    // The short form runs where its bound holds, the shifted form below
    // elsewhere within the exact-product bound; the exact-product form
    // runs in its place when it is switched off (JOB_SM_NO_SHIFT, a test
    // hook) or where it does not apply (index 0, a prime dividing b).
    //
    // Short form, for a segment where every lane has p (0.75 m + 1) < 2^53:
    // then |k p| < 2^53, so w = fma(k + M, -p, M p) = -k p is exact (M the
    // rounding constant, M p exact for p < 2^50), r m + w = r m - k p is
    // exact, and r' = (r m - k p) + c takes 4 instructions and the compare
    // (the same bound |r'| <= 0.75 p holds).
    //
    // Exact-product form, for a segment where every lane has
    // ((1/2 + 2^-30) p + n1) o < 2^53 with o the odd part of m (m = o 2^a).
    // Every step of every form leaves |r| <= (1/2 + 2^-30) p + n1 (the
    // reduced part, plus c < n1; the initial state is reduced), so
    // |r o| < 2^53: r o is exact and so is r m = (r o) 2^a (a power-of-two
    // scaling).  Then fma(-k, p, r m) has the exact value r m - k p,
    // |.| <= (1/2 + 2^-30) p < 2^53, so it is exact, and
    // r' = (r m - k p) + c takes 5 instructions and the compare.
    // Power-of-two bases (o = 1) always qualify.
template <int K>
__device__ void run_fp_sm(const JobArgs &J, const uint64_t (&p)[K], uint64_t ns, uint64_t ne, uint64_t idx,
                          uint64_t (&fin)[K])
{
    const uint64_t bb = J.param;
    Fmod F[K];
    double r[K], ph[K], pl[K], mp[K], pM[K], t[K];
    _Pragma("unroll")
    for (int q = 0; q < K; ++q) {
        F[q].init(p[q]);
        r[q] = sm_value_fp(F[q], bb, ns);
        record_hits(J, idx, p[q], r[q] == 0.0);
        const uint32_t L = 64 - __clzll(p[q]);
        const uint64_t phi = L > 32 ? p[q] & ~((1ull << (L - 32)) - 1) : p[q];
        ph[q] = (double)phi;
        pl[q] = (double)(p[q] - phi);
        pM[q] = F[q].p * Fmod::MAGIC;
        t[q] = 0.0;
    }
    uint32_t d;
    uint64_t nxt;
    digits_pow(bb, ns + 1, d, nxt);
    const double ob = (double)(bb >> __builtin_ctzll(bb)); // odd part of the base
    double m = 1.0, om = 1.0;                                // m = bb^d, om its odd part
    for (uint32_t i = 0; i < d; ++i) {
        m *= (double)bb;
        om *= ob;
    }
    _Pragma("unroll")
    for (int q = 0; q < K; ++q) mp[q] = m * F[q].invp;
    const double n1d = (double)J.n1; // c = n + 1 < n1 (n1 <= 2^32: JOB_SM_NARROW)
    double inv_om = 0.0, inv_om_of = 0.0; // 1/om, and the om it belongs to
    uint64_t n = ns; // index of the current state
    // The shifted form holds t = r - n instead of r (`shifted`); index n
    // itself is then not tested yet: its test is the next step's, or the
    // one that follows restoring r.
    bool shifted = false;
    while (n + 1 < ne) {
        // steps n -> n+1 .. seg_last -> seg_last+1 share m = b^digits(n+1)
        // (digits grow when n+1 reaches nxt); each form also keeps its
        // scalar sequence inside one binade
        uint64_t seg_last = nxt - 2 < ne - 2 ? nxt - 2 : ne - 2;
        // one wave-uniform choice per segment (m is fixed inside it)
        const bool free_form = !(J.flags & JOB_SM_FULL_STEP);
        // the short step, where it is exact, runs as fast as the shifted
        // one and needs no state change (measured: 1-2 % faster)
        bool short_form = free_form && !(J.flags & JOB_SM_NO_SHORT);
        if (short_form) {
            bool out = false;
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) out |= !(F[q].p * (0.75 * m + 1.0) < SM_SHORT_LIMIT);
            short_form = __ballot(out) == 0;
        }
        bool exact_ok = !short_form && free_form;
        if (exact_ok) {
            bool out = false;
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) out |= !((F[q].p * SM_EXACT_HALF + n1d) * om < SM_SHORT_LIMIT);
            exact_ok = __ballot(out) == 0;
        }
        bool shift_ok = exact_ok && n != 0 && !(J.flags & JOB_SM_NO_SHIFT);
        if (shift_ok) {
            bool small = false;
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) small |= p[q] <= bb;
            shift_ok = __ballot(small) == 0;
        }
        if (shift_ok) {
            // Shifted form: the state is t = r - n (n the last appended
            // number), so one fma forms r m = (t + n) m = t m + (n m) with the
            // wave-uniform n m from the scalar unit, exactly (the exact-product
            // bound: |r| o < 2^53 and n o < 2^53).  t' = r m - k p is exact
            // (|.| <= (1/2 + 2^-30) p) and equals S(n+1) - (n+1) mod p.  The
            // test of index n needs no appended number: S(n) = 0 mod p <=>
            // S(n) m = 0 <=> t' == 0 (m = b^d is prime to p > b).  So a step
            // is 4 instructions and the test, and the tests lag one index:
            // the first step after entering the form repeats the test of
            // index n (skipped), and the last index is tested when r = t + n
            // is restored.  n m is stepped as a bit pattern inside the
            // binade of n o (m = o 2^a), so a segment also ends where n o
            // leaves its binade; t carries over to the next segment.
            if (inv_om_of != om) { // 1/o, once per digit count
                inv_om = 1.0 / om;
                inv_om_of = om;
            }
            const uint64_t on = scalar_u64((uint64_t)om);
            const uint32_t e2 = 63 - __clzll(n * on);
            const uint64_t last_i = binade_last(e2, on, inv_om);
            if (last_i < seg_last) seg_last = last_i;
            const uint32_t cs = (uint32_t)(seg_last - n + 1);
            const double mv = opaque_vgpr(m);
            // n < n1 <= 2^32; the pattern is made scalar (m lives in a vector register)
            uint64_t cmb = scalar_u64(__double_as_longlong((double)(uint32_t)n * m));
            const uint64_t dcm = scalar_u64(on << (52 - e2));
            auto step = [&](double (&v)[K], uint64_t &s) {
                _Pragma("unroll")
                for (int q = 0; q < K; ++q) {
                    const double rm = fma(v[q], mv, __longlong_as_double(s));
                    const double k = fma(rm, F[q].invp, Fmod::MAGIC) - Fmod::MAGIC;
                    v[q] = fma(-k, F[q].p, rm);
                }
                s += dcm;
            };
            // entering the form: t = r - n, and the first step repeats the
            // test of index n (recorded already), so it runs untested
            const uint32_t done = shifted ? 0u : 1u;
            if (done) {
                _Pragma("unroll")
                for (int q = 0; q < K; ++q) t[q] = r[q] - (double)(uint32_t)n;
                step(t, cmb);
                shifted = true;
            }
            // (wave-uniform values; the compiler cannot prove it for `shifted`)
            cmb = scalar_u64(cmb);
            sm_blocks<K, true>(J, p, scalar_u64(idx + done), __builtin_amdgcn_readfirstlane(cs - done), t, cmb,
                               step);
            n = seg_last + 1;
            idx += cs;
        } else {
            if (shifted) { // back to r; index n is tested here
                _Pragma("unroll")
                for (int q = 0; q < K; ++q) {
                    r[q] = t[q] + (double)(uint32_t)n;
                    record_hits(J, idx, p[q], r[q] == 0.0);
                }
                shifted = false;
            }
            // the appended number c = n+1 is stepped inside its binade
            const uint64_t c0 = n + 1;
            const uint32_t e = 63 - __clzll(c0);
            if ((2ull << e) - 2 < seg_last) seg_last = (2ull << e) - 2;
            const uint32_t cnt = (uint32_t)(seg_last - n + 1); // below 2^32: n1 <= 2^32
            const uint64_t delta = 1ull << (52 - e);
            uint64_t cb = __double_as_longlong((double)c0);
            if (short_form) {
                sm_blocks<K, false>(J, p, idx + 1, cnt, r, cb, [&](double (&v)[K], uint64_t &s) {
                    _Pragma("unroll")
                    for (int q = 0; q < K; ++q) {
                        const double k1 = fma(v[q], mp[q], Fmod::MAGIC);
                        v[q] = fma(v[q], m, fma(k1, -F[q].p, pM[q])) + __longlong_as_double(s);
                    }
                    s += delta;
                });
            } else if (exact_ok) {
                sm_blocks<K, false>(J, p, idx + 1, cnt, r, cb, [&](double (&v)[K], uint64_t &s) {
                    _Pragma("unroll")
                    for (int q = 0; q < K; ++q) {
                        const double rm = v[q] * m;
                        const double k = fma(v[q], mp[q], Fmod::MAGIC) - Fmod::MAGIC;
                        v[q] = fma(-k, F[q].p, rm) + __longlong_as_double(s);
                    }
                    s += delta;
                });
            } else {
                sm_blocks<K, false>(J, p, idx + 1, cnt, r, cb, [&](double (&v)[K], uint64_t &s) {
                    _Pragma("unroll")
                    for (int q = 0; q < K; ++q) {
                        const double k = fma(v[q], mp[q], Fmod::MAGIC) - Fmod::MAGIC;
                        const double u = fma(v[q], m, -(k * ph[q]));
                        v[q] = fma(-k, pl[q], u + __longlong_as_double(s));
                    }
                    s += delta;
                });
            }
            idx += cnt;
            n = seg_last + 1;
        }
        if (n + 1 == nxt) { // the next appended number has one more digit
            m *= (double)bb;
            om *= ob;
            _Pragma("unroll")
            for (int q = 0; q < K; ++q) mp[q] = m * F[q].invp;
            nxt = sat_mul(nxt, bb);
        }
    }
    _Pragma("unroll")
    for (int q = 0; q < K; ++q) {
        if (shifted) {
            r[q] = t[q] + (double)(uint32_t)n;
            record_hits(J, idx, p[q], r[q] == 0.0);
        }
        fin[q] = F[q].canon(r[q], p[q]);
    }
}

// The FP64 form of Fast<FAM>::run for a wave whose primes are all below
// FP_PMAX (and, for smarandache, above 8 n1).  Returns the plain
// residue S(ne-1) mod p for the test hook.
template <uint32_t FAM>
__device__ uint64_t run_fp(const JobArgs &J, uint64_t p, uint64_t ns, uint64_t ne, uint64_t idx, uint64_t steps)
{
    Fmod F;
    F.init(p);
    if constexpr (FAM == FAM_MERSENNE || FAM == FAM_PROTH || FAM == FAM_RIESEL) {
        // x = c 2^n, hit when x == t (t = -1 for proth, +1 otherwise)
        double x = F.pow2(ns);
        if constexpr (FAM != FAM_MERSENNE) x = F.mul(F.of(J.param < p ? J.param : J.param % p), x);
        x = F.red(x);
        const double tgt = (FAM == FAM_PROTH) ? -1.0 : 1.0;
        record_hits(J, idx, p, x == tgt);
        uint32_t left = (uint32_t)steps;
        {
            // targets t 2^-(j+1), j < DOUBLING_WORDS, canonical, as packed
            // words (one register each; 128 words leave 3 waves per SIMD,
            // which the loop runs at full rate)
            uint64_t Tw[DOUBLING_WORDS / 2];
            {
                double t = tgt;
                _Pragma("unroll")
                for (int j = 0; j < DOUBLING_WORDS; j += 2) {
                    t = half_canon(F, t);
                    const uint64_t w0 = low_word(t);
                    t = half_canon(F, t);
                    const uint64_t w1 = low_word(t);
                    uint64_t pair = w0 | (w1 << 32);
                    asm volatile("" : "+v"(pair)); // keep the words packed in one register pair
                    Tw[j / 2] = pair;
                }
            }
            // 2^16 mod p by one exact reduction, 2^64 and 2^128 mod p by squarings
            const double C16 = F.red(65536.0), C32 = mulc_canon(F, C16, C16), C64 = mulc_canon(F, C32, C32);
            x = doubling_words<DOUBLING_WORDS>(J, p, idx, left / DOUBLING_WORDS, x, Tw, F,
                                               mulc_canon(F, C64, C64), tgt);
            left %= DOUBLING_WORDS;
            x = doubling_words<64>(J, p, idx, left / 64, x, Tw, F, C64, tgt);
            left %= 64;
            x = doubling_words<16>(J, p, idx, left / 16, x, Tw, F, C16, tgt);
            left %= 16;
        }
        for (; left; --left) {
            x = F.red(x + x);
            ++idx;
            const uint64_t m = __ballot(x == tgt);
            if (m) record_hits_mask(J, idx, p, (m >> __lane_id()) & 1, m);
        }
        return F.canon(FAM == FAM_PROTH ? x + 1.0 : x - 1.0, p);
    } else if constexpr (FAM == FAM_PROTH_FIXED_N || FAM == FAM_RIESEL_FIXED_N) {
        // x = S(k) = k 2^n +- 1; step x += c with c = 2^n; hit x == 0
        const double c = F.red(F.pow2(J.param));
        const uint64_t k = ns < (1ull << 52) ? ns : ns % p;
        double x = F.red(F.mul(F.of(k), c) + (FAM == FAM_PROTH_FIXED_N ? 1.0 : -1.0));
        record_hits(J, idx, p, x == 0.0);
        // T[j] = -(j+1) c (j < 64), reduced and kept as the low 32-bit words
        // of T + M (M the rounding constant), two per register pair: a block
        // of 64 indices costs the word of x, 64 compares and the advance by
        // 64 c; a block with a word match is replayed exactly
        uint64_t Tw[32];
        {
            double t = 0.0;
            _Pragma("unroll")
            for (int j = 0; j < 64; j += 2) {
                t = F.red(t - c);
                const uint64_t w0 = low_word(t);
                t = F.red(t - c);
                const uint64_t w1 = low_word(t);
                uint64_t pair = w0 | (w1 << 32);
                asm volatile("" : "+v"(pair)); // keep the words packed in one register pair
                Tw[j / 2] = pair;
            }
        }
        const double c16 = F.red(16.0 * c), c64 = F.red(4.0 * c16);
        uint32_t left = (uint32_t)steps;
        x = fixedn_blocks<64>(J, p, idx, left / 64, x, Tw, F, c, c64);
        left %= 64;
        x = fixedn_blocks<16>(J, p, idx, left / 16, x, Tw, F, c, c16);
        left %= 16;
        for (; left; --left) {
            x = F.red(x + c);
            ++idx;
            const uint64_t m = __ballot(x == 0.0);
            if (m) record_hits_mask(J, idx, p, (m >> __lane_id()) & 1, m);
        }
        return F.canon(x, p);
    } else if constexpr (FAM == FAM_FIBONACCI || FAM == FAM_LUCAS) {
        if (J.n0 != 0 && steps >= THETA_MIN_STEPS && __ballot(p == 5) == 0) {
            // theta stride form: x = theta^n, hit when x == eps (eps = 1 for
            // Fibonacci, -1 for Lucas).  The THETA_STRIDE<FAM> indices after x_n
            // are hits when x_n == eps theta^-j.  Every power of theta has
            // norm u^2 + uv - v^2 = 1, so for a given v at most two u qualify
            // (u and -v - u): the loop compares (a 32-bit word of) the v
            // coordinate only, one compare per index, and a match is
            // confirmed on both coordinates before it is recorded.  x advances THETA_STRIDE<FAM> indices by one
            // ring product with theta^THETA_STRIDE<FAM>.
            const double eps = FAM == FAM_FIBONACCI ? 1.0 : -1.0;
            Rphi x = rphi_theta_pow(F, ns);
            record_hits(J, idx, p, x.u == eps && x.v == 0.0);
            // targets: the low 32 bits of v + M (M the rounding constant, which
            // moves an integer |v| < 2^51 into the low bits of the pattern).
            // Equal v give equal words; a word match that is not a hit (values
            // 2^32 apart) is rejected by the exact replay.  Two words share a
            // 64-bit register pair, so the targets take THETA_STRIDE<FAM> registers.
            // The v coordinates of eps theta^-j follow v_(j+1) = -3 v_j - v_(j-1)
            // from v_0 = 0, v_1 = eps: theta^-1 = phi - 2 is a root of
            // y^2 + 3y + 1 (theta + theta^-1 = -(phi^2 + psi^2) = -L(2) = -3),
            // so every coordinate of eps theta^-j obeys that recurrence.  The
            // sum stays below 2p + 1, so red gives the same (symmetric)
            // representative as the ring product does for x.
            uint64_t Tw[THETA_STRIDE<FAM> / 2];
            {
                double va = 0.0, vb = eps; // (v_j, v_(j+1))
                _Pragma("unroll")
                for (int j = 0; j < THETA_STRIDE<FAM>; j += 2) {
                    const uint64_t w0 = low_word(vb);
                    va = F.red(fma(-3.0, vb, -va));
                    const uint64_t w1 = low_word(va);
                    vb = F.red(fma(-3.0, va, -vb));
                    uint64_t pair = w0 | (w1 << 32);
                    asm volatile("" : "+v"(pair)); // keep the words packed in one register pair
                    Tw[j / 2] = pair;
                }
            }
            const Rphi g = rphi_theta_pow(F, THETA_STRIDE<FAM>);
            const double gw = F.red(g.u + g.v);
            const uint32_t steps32 = (uint32_t)steps;
            uint32_t k = 0;
            for (; k + THETA_STRIDE<FAM> <= steps32; k += THETA_STRIDE<FAM>) {
                const uint32_t xw = low_word(x.v);
                uint64_t any = 0;
                _Pragma("unroll")
                for (int q = 0; q < THETA_STRIDE<FAM> / 16; ++q) any |= words_match16(xw, Tw + 8 * q);
                if (any) {
                    // replay the block: x theta^(j+1) == eps on both coordinates
                    // (x is made opaque so that the loop's masks are not kept
                    // live for this rare path)
                    const uint64_t at = hit_base(idx, k);
                    Rphi y = {opaque_vgpr(x.u), opaque_vgpr(x.v)};
                    for (int j = 0; j < THETA_STRIDE<FAM>; ++j) {
                        y = rphi_mul_theta(F, y);
                        record_hits(J, at + 1 + j, p, y.u == eps && y.v == 0.0);
                    }
                }
                x = rphi_mul_const(F, x, g.u, g.v, gw);
            }
            idx += k;
            for (; k < steps32; ++k) {
                x = rphi_mul_theta(F, x);
                ++idx;
                record_hits(J, idx, p, x.u == eps && x.v == 0.0);
            }
            // S(ne-1), by the doubling formulas, only for the test hook that
            // asks for final residues
            if (!J.final_res) return 0;
            double fa, fb;
            fib_pair_fp(F, ne - 1, fa, fb); // (F(ne-1), F(ne))
            if constexpr (FAM == FAM_FIBONACCI) return F.canon(fa, p);
            else return F.canon(F.red(fb + fb - fa), p); // L(n) = 2F(n+1) - F(n)
        }
        double a = 0.0, b = 0.0; // (S(n-1), S(n)); n0 = 0 keeps {0, 0} (fibonacci.hpp:68-71)
        if (J.n0 != 0) {
            double fa, fb;
            fib_pair_fp(F, ns - 1, fa, fb);
            if constexpr (FAM == FAM_FIBONACCI) {
                a = fa;
                b = fb;
            } else { // L(n-1) = 2F(n) - F(n-1), L(n) = F(n) + 2F(n-1)
                a = F.red(fb + fb - fa);
                b = F.red(fb + fa + fa);
            }
        }
        record_hits(J, idx, p, b == 0.0);
        CTD_UNROLLED_LOOP_MASKS8(
            {
                _Pragma("unroll")
                for (int j = 0; j < 8; j += 2) {
                    a = F.red(a + b);
                    mk[j] = __ballot(a == 0.0);
                    b = F.red(a + b);
                    mk[j + 1] = __ballot(b == 0.0);
                }
            },
            { const double t = F.red(a + b); a = b; b = t; }, __ballot(b == 0.0))
        return F.canon(b, p);
    } else if constexpr (FAM == FAM_TRIBONACCI) {
        double a, b, c;
        trib_triple_fp(F, ns, a, b, c);
        record_hits(J, idx, p, a == 0.0);
        // lazy form (trib_lazy_block): the longest block whose terms stay
        // exact for every prime of the wave (0.75 S p < 2^53 with S the sum
        // of the last term's coefficients)
        if (__ballot(p >= (1ull << 43)) == 0) trib_lazy_loop<11>(J, F, p, a, b, c, idx, steps);
        else if (__ballot(p >= (1ull << 44)) == 0) trib_lazy_loop<10>(J, F, p, a, b, c, idx, steps);
        else if (__ballot(p >= (1ull << 45)) == 0) trib_lazy_loop<8>(J, F, p, a, b, c, idx, steps);
        else if (__ballot(p >= (1ull << 46)) == 0) trib_lazy_loop<7>(J, F, p, a, b, c, idx, steps);
        else if (__ballot(p >= (1ull << 47)) == 0) trib_lazy_loop<6>(J, F, p, a, b, c, idx, steps);
        else trib_lazy_loop<5>(J, F, p, a, b, c, idx, steps); // p < FP_PMAX = 2^48
        return F.canon(a, p);
    } else if constexpr (FAM == FAM_SMARANDACHE) {
        const uint64_t pp[1] = {p};
        uint64_t fin[1];
        run_fp_sm<1>(J, pp, ns, ne, idx, fin);
        return fin[0];
    } else {
        return 0;
    }
}

} // namespace ctd
