"""CPU checks of the exactness arguments of the FP64 residue forms
(cutrialdive_amd/csrc/ctd_tf_fp.hpp), with every rounding emulated exactly.

Python's float(Fraction) rounds to nearest (ties to even), as the hardware
does, so fma(a, b, c) = float(Fraction(a) * b + c) is the fused multiply-add
with its single rounding.

The Tribonacci lazy form (trib_lazy_block) forms up to 11 unreduced terms
x = T(n+3+j) per block and tests most of them with a filter: the low 32-bit
word of y = fma(x, 1/p, 1.5 2^20) is 0 for every multiple of p, and a block
whose filter fires tests those terms again with the exact reduction.  These
tests pin both halves (no hit is missed; a near miss passes the filter and
is rejected by the exact test) and the growth bound that keeps the terms
exact.
"""
from __future__ import annotations

import random
import struct
from fractions import Fraction

import pytest

BIAS = 1572864.0  # 1.5 * 2^20
MAGIC = 6755399441055744.0  # 1.5 * 2^52


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def low_word(y: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", y))[0] & 0xFFFFFFFF


def filter_passes(x: float, invp: float) -> bool:
    return low_word(fma(x, invp, BIAS)) == 0


def neg_round(x: float, p: float, invp: float) -> float:
    # trib_neg_round: fma(fma(x, 1/p, M), -p, M p) = -round(x / p) p
    return fma(fma(x, invp, MAGIC), -p, p * MAGIC)


def coefficients(L: int):
    """(A, B, C) with x[j] = A a + B b + C c for the block's terms."""
    a, b, c = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    add = lambda u, v: tuple(s + t for s, t in zip(u, v))
    sc = lambda k, u: tuple(k * s for s in u)
    x = [add(a, add(b, c))]
    x.append(add(sc(2, x[0]), sc(-1, a)))
    x.append(add(sc(2, x[1]), sc(-1, b)))
    x.append(add(sc(2, x[2]), sc(-1, c)))
    for j in range(4, L):
        x.append(add(sc(2, x[j - 1]), sc(-1, x[j - 4])))
    return x


def block(a: float, b: float, c: float, p: int, L: int):
    """The block as the kernel computes it: the terms, the masks of the
    indices n+1 .. n+L after the filter and after the exact re-test, the
    filter's own masks, and the next state."""
    pf, invp = float(p), 1.0 / float(p)
    x = [a + (b + c)]
    x.append(fma(2.0, x[0], -a))
    x.append(fma(2.0, x[1], -b))
    x.append(fma(2.0, x[2], -c))
    for j in range(4, L):
        x.append(fma(2.0, x[j - 1], -x[j - 4]))
    filt = [filter_passes(x[j], invp) for j in range(L - 3)]
    w = [neg_round(x[L - 3 + j], pf, invp) for j in range(3)]
    exact = [x[j] == -neg_round(x[j], pf, invp) if filt[j] else False for j in range(L - 3)]
    masks = [b == 0.0, c == 0.0] + exact + [x[L - 3] == -w[0]]
    state = (x[L - 3] + w[0], x[L - 2] + w[1], x[L - 1] + w[2])
    return x, masks, filt, state


def primes_near(limit: int, count: int, rng):
    import sympy

    out = []
    p = limit
    while len(out) < count:
        p = int(sympy.prevprime(p))
        out.append(p)
    out += [int(sympy.nextprime(rng.randrange(1 << 31, limit))) for _ in range(count)]
    return out


# the block length the kernel takes below each prime limit (run_fp, FAM_TRIBONACCI)
BLOCKS = [(1 << 43, 11), (1 << 44, 10), (1 << 45, 8), (1 << 46, 7), (1 << 47, 6), (1 << 48, 5)]


def test_growth_bound_keeps_terms_exact():
    # |x[j]| <= (A + B + C) 0.75 p, the sums 3, 5, 9, ..., 1201: the last
    # term of each block length stays below 2^53 below its prime limit, the
    # next longer block would not, and |x / p| stays below 2^19, where y
    # keeps 32 fraction bits
    sums = [sum(t) for t in coefficients(11)]
    assert sums == [3, 5, 9, 17, 31, 57, 105, 193, 355, 653, 1201]
    assert all(min(t) >= 0 for t in coefficients(11))
    for limit, L in BLOCKS:
        assert 0.75 * sums[L - 1] * limit < 2 ** 53, (limit, L)
        if L < 11:
            assert 0.75 * sums[L] * limit > 2 ** 53, (limit, L + 1)
    assert 0.75 * sums[10] < 2 ** 19


@pytest.mark.parametrize("limit,L", BLOCKS)
def test_filter_passes_every_multiple(limit, L):
    rng = random.Random(limit)
    bound = int(0.75 * sum(coefficients(L)[L - 1])) + 1
    for p in primes_near(limit, 24, rng) + [3, 5, 7, 65537, (1 << 31) - 1]:
        invp = 1.0 / float(p)
        for i in list(range(-bound, -bound + 40)) + list(range(-20, 21)) + list(range(bound - 40, bound + 1)):
            x = float(i * p)
            assert abs(i * p) < 2 ** 53
            assert filter_passes(x, invp), (p, i)
            assert x == -neg_round(x, float(p), invp)


def test_filter_near_misses_and_small_primes():
    rng = random.Random(7)
    # below 2^32 a term that is not a multiple never passes
    for p in (3, 5, 7, 11, 65537, 4294967291):
        invp = 1.0 / p
        for x in range(-3000, 3000):
            assert filter_passes(float(x), invp) == (x % p == 0), (p, x)
    # large primes: x = i p + r with 0 < |r| < p 2^-34 passes the filter (a
    # near miss) and fails the exact test; |r| > p 2^-32 never passes
    for p in primes_near(1 << 43, 16, rng):
        invp, pf = 1.0 / p, float(p)
        for i in (-900, -1, 0, 1, 450, 900):
            for r in (1, -1, 3, -7, 100, 5000):
                x = float(i * p + r)
                if abs(r) << 34 < p:
                    assert filter_passes(x, invp), (p, i, r)
                elif abs(r) << 32 > p:
                    assert not filter_passes(x, invp), (p, i, r)
                assert x != -neg_round(x, pf, invp)
            for r in (p // 2, -(p // 3), p // 1000):
                assert not filter_passes(float(i * p + r), invp)


@pytest.mark.parametrize("limit,L", BLOCKS)
def test_block_masks_equal_exact_divisibility(limit, L):
    """Random and planted states: after the exact re-test every mask equals
    'the true integer term is divisible by p'; the filter masks contain them;
    the next state is congruent to the true terms and reduced."""
    rng = random.Random(L)
    co = coefficients(L)
    for p in primes_near(limit, 12, rng) + [101, 3, 65521]:
        lim = int(0.75 * p)
        for trial in range(40):
            a, b = rng.randint(-lim, lim), rng.randint(-lim, lim)
            j = rng.randrange(L)
            A, B, C = co[j]
            if trial % 2 and C % p:
                # plant a zero at term j: solve A a + B b + C c = 0 (mod p) for c
                c = (-(A * a + B * b) * pow(C, -1, p)) % p
                if c > lim:
                    c -= p
            else:
                c = rng.randint(-lim, lim)
            x, masks, filt, state = block(float(a), float(b), float(c), p, L)
            true = [A * a + B * b + C * c for A, B, C in co]
            assert [int(v) for v in x] == true
            assert all(abs(v) < 2 ** 53 for v in true)
            want = [b % p == 0, c % p == 0] + [t % p == 0 for t in true[: L - 2]]
            assert masks == want, (p, a, b, c)
            assert all(f or t % p for f, t in zip(filt, true)), (p, a, b, c)
            for s, t in zip(state, true[L - 3 :]):
                assert (int(s) - t) % p == 0
                assert abs(s) <= p / 2 + 2


# ---------------------------------------------------------------- Smarandache steps
#
# run_fp_sm (ctd_tf_fp.hpp) steps r_{n+1} = r_n m + (n+1) mod p with
# m = b^digits(n+1) < 2^21, c = n+1 < n1 <= 2^32 and p > 8 n1, p < 2^48, in
# one of four forms chosen per segment by the bounds below.  Each emulation
# checks that every operation the kernel relies on is exact (its rounded
# result equals the exact rational value), that the new state is congruent
# to the true residue, and that it stays inside the bound the next step
# assumes.

SM_SHORT_LIMIT = 9007199254740992.0 - 1073741824.0  # 2^53 - 2^30
SM_EXACT_HALF = 0.5000001


def exact(value, want):
    assert Fraction(value) == want, (value, want)
    return value


def sm_short(v, m, c, p):
    """Short step (p (0.75 m + 1) < 2^53 - 2^30): k1 = fma(v, m/p, M),
    v' = fma(v, m, fma(k1, -p, M p)) + c."""
    pf, mf = float(p), float(m)
    mp = mf * (1.0 / pf)
    pM = exact(pf * MAGIC, Fraction(p) * Fraction(MAGIC))
    k1 = fma(v, mp, MAGIC)
    w = exact(fma(k1, -pf, pM), -(Fraction(k1) - Fraction(MAGIC)) * p)
    x = exact(fma(v, mf, w), Fraction(v) * m + Fraction(w))
    return exact(x + c, Fraction(x) + c)


def sm_exact_product(v, m, c, p):
    """Exact-product step (((1/2 + 2^-30) p + n1) o < 2^53, o the odd part
    of m): r m is exact, v' = fma(-k, p, r m) + c."""
    pf, mf = float(p), float(m)
    mp = mf * (1.0 / pf)
    rm = exact(v * mf, Fraction(v) * m)
    k = fma(v, mp, MAGIC) - MAGIC
    x = exact(fma(-k, pf, rm), Fraction(rm) - Fraction(k) * p)
    return exact(x + c, Fraction(x) + c)


def sm_general(v, m, c, p):
    """General step: k = round(v m / p), u = fma(v, m, -(k p_hi)) exact with
    p_hi the top 32 significant bits of p, v' = fma(-k, p_lo, u + c)."""
    pf, mf = float(p), float(m)
    mp = mf * (1.0 / pf)
    L = p.bit_length()
    ph = p & ~((1 << (L - 32)) - 1) if L > 32 else p
    pl = p - ph
    k = fma(v, mp, MAGIC) - MAGIC
    kph = exact(k * float(ph), Fraction(k) * ph)
    u = exact(fma(v, mf, -kph), Fraction(v) * m - Fraction(kph))
    uc = exact(u + c, Fraction(u) + c)
    return exact(fma(-k, float(pl), uc), Fraction(uc) - Fraction(k) * pl)


def sm_shifted(t, m, n, p):
    """Shifted step on t = r - n: rm = fma(t, m, n m) exact, t' = rm - k p =
    S(n+1) - (n+1) mod p; t' == 0 tests index n."""
    pf, mf = float(p), float(m)
    nm = exact(float(n) * mf, Fraction(n) * m)
    rm = exact(fma(t, mf, nm), Fraction(t) * m + Fraction(nm))
    k = fma(rm, 1.0 / pf, MAGIC) - MAGIC
    return exact(fma(-k, pf, rm), Fraction(rm) - Fraction(k) * p)


def digits(n, b):
    d = 0
    while n:
        n //= b
        d += 1
    return d


def sm_residue(b, n, p):
    r = 0
    for i in range(1, n + 1):
        r = (r * b ** digits(i, b) + i) % p
    return r


def odd_part(m):
    while m % 2 == 0:
        m //= 2
    return m


def run_sm_form(form, b, n0, n1, p):
    """Steps n0 -> n1 - 1 in one form from S(n0) mod p; every state is
    checked against the true residue and its bound."""
    r = sm_residue(b, n0, p)
    v = float(r if r <= p // 2 else r - p)
    shifted = form is sm_shifted
    if shifted:
        v = v - n0  # t = r - n
    zeros = 0
    for n in range(n0, n1 - 1):
        m = b ** digits(n + 1, b)
        true_next = (r * m + n + 1) % p
        if shifted:
            t2 = form(v, m, n, p)
            # t2 = S(n+1) - (n+1) and it is 0 exactly when S(n) is
            assert (int(t2) - (true_next - (n + 1))) % p == 0
            assert (t2 == 0.0) == (r == 0)
            assert abs(t2) <= SM_EXACT_HALF * p
            v = t2
        else:
            v2 = form(v, m, n + 1, p)
            assert (int(v2) - true_next) % p == 0
            bound = 0.75 * p if form is not sm_exact_product else SM_EXACT_HALF * p + n1
            assert abs(v2) <= bound
            assert (v2 == 0.0) == (true_next == 0)
            v = v2
        zeros += r == 0
        r = true_next
    return zeros


def segment(b, m, n0, n1, length):
    """[lo, hi): steps n -> n+1 inside [n0, n1) that share the multiplier m
    (the kernel picks a form per such segment), at most `length` of them."""
    ns = [n for n in range(n0, n1 - 1) if b ** digits(n + 1, b) == m]
    return ns[0], min(ns[-1] + 2, ns[0] + length + 1)


def largest_prime_below(x):
    import sympy

    return int(sympy.prevprime(int(x)))


@pytest.mark.parametrize("b", [10, 13, 3, 16, 7])
def test_smarandache_short_step_at_its_limit(b):
    # the largest prime the short step admits for each multiplier of [n0, n1)
    n0, n1 = 90, 1200
    for m in sorted({b ** digits(n, b) for n in range(n0 + 1, n1)}):
        p = largest_prime_below(SM_SHORT_LIMIT / (0.75 * m + 1.0))
        if p < 8 * n1 or p >= 1 << 48:
            continue
        assert p * (0.75 * m + 1.0) < SM_SHORT_LIMIT
        lo, hi = segment(b, m, n0, n1, 300)
        run_sm_form(sm_short, b, lo, hi, p)


@pytest.mark.parametrize("b", [10, 6, 12, 4, 2])
def test_smarandache_exact_product_and_shifted_steps_at_their_limit(b):
    n0, n1 = 90, 1200
    for m in sorted({b ** digits(n, b) for n in range(n0 + 1, n1)}):
        o = odd_part(m)
        p = largest_prime_below((SM_SHORT_LIMIT / o - n1) / SM_EXACT_HALF)
        if p >= 1 << 48:
            p = largest_prime_below(1 << 48)
        assert (p * SM_EXACT_HALF + n1) * o < SM_SHORT_LIMIT
        lo, hi = segment(b, m, n0, n1, 200)
        run_sm_form(sm_exact_product, b, lo, hi, p)
        run_sm_form(sm_shifted, b, max(lo, 1), hi, p)


@pytest.mark.parametrize("b", [13, 15, 9, 10, 11])
def test_smarandache_general_step_large_multipliers(b):
    # the largest narrow multipliers (up to 13^5, 15^5) with primes up to 2^48
    rng = random.Random(b)
    for n0 in (b ** 3 - 3, b ** 4 - 3, 100003):
        m = b ** digits(n0 + 2, b)
        if m >= 1 << 21:
            continue
        n1 = n0 + 150
        for p in (largest_prime_below(1 << 48), largest_prime_below(1 << 43),
                  int(__import__("sympy").nextprime(rng.randrange(max(8 * n1, 1 << 38), 1 << 47)))):
            run_sm_form(sm_general, b, n0, n1, p)


def test_smarandache_steps_find_planted_zeros():
    # small primes above 8 n1, where zeros are frequent: every form (within
    # its bound) reports exactly the true zeros
    import sympy

    b, n0, n1 = 10, 20, 700
    total = 0
    for p in sympy.primerange(8 * n1 + 1, 8 * n1 + 400):
        for form in (sm_short, sm_exact_product, sm_general, sm_shifted):
            total += run_sm_form(form, b, n0, n1, p)
    assert total > 0


# ---------------------------------------------------------------- doubling stride arithmetic
#
# The doubling families' word stride (ctd_tf_fp.hpp) keeps x = k 2^n mod p
# canonical (|x| <= (p-1)/2), so that equal residues have equal doubles and
# their low words can be compared against precomputed targets t 2^-j.  The
# targets come from half_canon, the state advances by mulc_canon with
# C = 2^128 mod p.


def quot(v, p):
    return fma(v, 1.0 / float(p), MAGIC) - MAGIC


def red(v, p):
    return fma(-quot(v, p), float(p), v)


def mulc_canon(x, C, p):
    hi = x * C
    lo = exact(fma(x, C, -hi), Fraction(x) * Fraction(C) - Fraction(hi))
    t = exact(fma(-quot(hi, p), float(p), hi), Fraction(hi) - Fraction(quot(hi, p)) * p)
    return red(exact(t + lo, Fraction(t) + Fraction(lo)), p)


def half_canon(t, p):
    import math

    h = 0.5 * t
    f = h - math.floor(h)
    return fma(-f, math.copysign(float(p), t), h)


def centered(r, p):
    r %= p
    return r - p if r > p // 2 else r


@pytest.mark.parametrize("limit", [1 << 36, 1 << 44, 1 << 48])
def test_mulc_canon_and_half_canon_are_exact_and_canonical(limit):
    rng = random.Random(limit)
    for p in primes_near(limit, 10, rng)[:14]:
        h = (p - 1) // 2
        cases = [(h, h), (-h, h), (h, -h), (1, h), (h - 1, 2), (-1, -1)]
        cases += [(rng.randint(-h, h), rng.randint(-h, h)) for _ in range(40)]
        for x, C in cases:
            got = mulc_canon(float(x), float(C), p)
            assert got == centered(x * C, p), (p, x, C)
            assert abs(got) <= h
        for t in [h, -h, 1, -1, 0, 2] + [rng.randint(-h, h) for _ in range(40)]:
            got = half_canon(float(t), p)
            assert got == centered(t * pow(2, -1, p), p), (p, t)


# ---------------------------------------------------------------- Fmod (ctd_tf_fp.hpp)
#
# Fmod::mul(a, b) = fma(-round(hi/p), p, hi) + lo with hi = a b rounded and
# lo = fma(a, b, -hi) its exact error (Dekker): every operation is exact and
# the result is congruent to a b; Fmod::red(v) = fma(-round(v/p), p, v).  The
# set-up code (powers, doubling formulas, jump-ahead) chains them on values
# of size about p, and its results feed the loops after one more red().


def fmod_mul(a, b, p):
    pf = float(p)
    hi = a * b
    lo = exact(fma(a, b, -hi), Fraction(a) * Fraction(b) - Fraction(hi))
    q = quot(hi, p)
    t = exact(fma(-q, pf, hi), Fraction(hi) - Fraction(q) * p)
    return exact(t + lo, Fraction(t) + Fraction(lo))


@pytest.mark.parametrize("limit", [1 << 32, 1 << 44, 1 << 48])
def test_fmod_mul_and_red_exact_over_chains(limit):
    rng = random.Random(limit + 1)
    for p in primes_near(limit, 6, rng):
        x = float(rng.randint(-(p // 2), p // 2))
        y = float(rng.randint(-(p // 2), p // 2))
        ex, ey = int(x), int(y)
        # chains of products as the set-up runs them (squarings, products of
        # unreduced results, a reduction now and then): congruent, exact, and
        # never larger than 2 p in size
        for step in range(200):
            if step % 3 == 0:
                x, ex = fmod_mul(x, x, p), ex * ex % p
            elif step % 3 == 1:
                x, ex = fmod_mul(x, y, p), ex * ey % p
            else:
                x, ex = red(x + x, p), 2 * ex % p
            assert (int(x) - ex) % p == 0
            assert abs(x) <= 2 * p
        r = red(x, p)
        assert (int(r) - ex) % p == 0 and abs(r) <= 0.75 * p


# ---------------------------------------------------------------- theta form (Fibonacci / Lucas)
#
# The theta stride form works in Z_p[phi] / (phi^2 - phi - 1) with
# theta = -(1 + phi): F(n) = 0 mod p <=> theta^n = 1 and L(n) = 0 mod p <=>
# theta^n = -1 for p != 2, 5; the targets eps theta^-j have v coordinates
# v_(j+1) = -3 v_j - v_(j-1) from v_0 = 0, v_1 = eps; every power of theta
# has norm u^2 + u v - v^2 = 1, so a v coordinate leaves two candidates u.
# Exact integer arithmetic mod p, independent of the kernel's FP64 forms.


def rphi_mul(x, y, p):
    (xu, xv), (yu, yv) = x, y
    a, b = xu * yu, xv * yv
    c = (xu + xv) * (yu + yv)
    return ((a + b) % p, (c - a) % p)


def rphi_pow(x, e, p):
    r = (1, 0)
    while e:
        if e & 1:
            r = rphi_mul(r, x, p)
        x = rphi_mul(x, x, p)
        e >>= 1
    return r


def fib_lucas(n, p):
    """(F(n), L(n)) mod p by fast doubling."""
    def fd(k):
        if k == 0:
            return (0, 1)
        a, b = fd(k >> 1)
        c = a * ((2 * b - a) % p) % p
        d = (a * a + b * b) % p
        return (d, (c + d) % p) if k & 1 else (c, d)
    f, f1 = fd(n)
    return f, (2 * f1 - f) % p


def test_theta_form_algebra():
    import sympy

    rng = random.Random(5)
    theta = lambda p: ((-1) % p, (-1) % p)
    primes = [3, 7, 11, 13, 29, 89, 1009] + [int(sympy.nextprime(rng.randrange(1 << 30, 1 << 47))) for _ in range(12)]
    for p in primes:
        t = theta(p)
        assert rphi_mul(t, t, p) == rphi_mul(t, t, p)
        # the product formula against the definition with phi^2 = phi + 1
        for _ in range(10):
            x = (rng.randrange(p), rng.randrange(p))
            y = (rng.randrange(p), rng.randrange(p))
            want = ((x[0] * y[0] + x[1] * y[1]) % p, (x[0] * y[1] + x[1] * y[0] + x[1] * y[1]) % p)
            assert rphi_mul(x, y, p) == want
        # hits: F(n) = 0 <=> theta^n = 1, L(n) = 0 <=> theta^n = -1, and norm 1
        ns = list(range(0, 200)) + [rng.randrange(1, 1 << 40) for _ in range(30)]
        if p < 2000:
            # the rank of apparition and its multiples, where hits are
            a = next(n for n in range(1, 4 * p) if fib_lucas(n, p)[0] == 0)
            ns += [a, 2 * a, a // 2, 3 * a, a + 1]
        for n in ns:
            x = rphi_pow(t, n, p)
            f, l = fib_lucas(n, p)
            assert (f == 0) == (x == (1, 0)), (p, n)
            assert (l == 0) == (x == (p - 1, 0)), (p, n)
            u, v = x
            assert (u * u + u * v - v * v) % p == 1
        # targets: the v coordinates of eps theta^-j follow the recurrence
        tinv = rphi_pow(t, p * p - 2, p)  # theta^-1 (theta^(p^2-1) = 1 in the ring's unit group)
        assert rphi_mul(t, tinv, p) == (1, 0)
        for eps in (1, p - 1):
            va, vb = 0, eps % p
            y = (eps % p, 0)
            for j in range(1, 130):
                y = rphi_mul(y, tinv, p)
                assert y[1] == vb, (p, eps, j)
                va, vb = vb, (-3 * vb - va) % p
        # a v coordinate leaves two candidates for u: u and -v - u
        for _ in range(20):
            n = rng.randrange(1, 1 << 30)
            u, v = rphi_pow(t, n, p)
            cands = [w for w in ((u, v), ((-v - u) % p, v)) if (w[0] * w[0] + w[0] * v - v * v) % p == 1]
            assert (u, v) in cands


def test_theta_hits_at_known_large_factors():
    # q | F(m) for primes q in [2^37, 2^47) (tests/golden/fib_large_factors.json):
    # theta^m = 1 in Z_q[phi], and theta^(m/2) = -1 when L(m/2) = 0 as well
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fib_large_factors.json")
    with open(path) as f:
        pairs = [tuple(x) for x in json.load(f)["pairs"]]
    assert pairs
    for m, q in pairs:
        t = ((-1) % q, (-1) % q)
        assert fib_lucas(m, q)[0] == 0
        assert rphi_pow(t, m, q) == (1, 0)
        assert rphi_pow(t, m + 1, q) != (1, 0)
        if m % 2 == 0 and fib_lucas(m // 2, q)[1] == 0:
            assert rphi_pow(t, m // 2, q) == (q - 1, 0)
