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


def test_growth_bound_keeps_terms_exact():
    # |x[j]| <= (A + B + C) 0.75 p, the sums 3, 5, 9, ..., 1201; the last
    # term of an 11-index block stays below 2^53 for p < 2^43, of an 8-index
    # block for p < 2^44, and |x / p| stays below 2^19, where y keeps 32
    # fraction bits
    sums = [sum(t) for t in coefficients(11)]
    assert sums == [3, 5, 9, 17, 31, 57, 105, 193, 355, 653, 1201]
    assert all(min(t) >= 0 for t in coefficients(11))
    assert 0.75 * sums[10] * 2 ** 43 < 2 ** 53
    assert 0.75 * sums[7] * 2 ** 44 < 2 ** 53
    assert 0.75 * sums[10] * 2 ** 44 > 2 ** 53  # an 11-index block would not be exact at 2^44
    assert 0.75 * sums[10] < 2 ** 19


@pytest.mark.parametrize("limit,L", [(1 << 43, 11), (1 << 44, 8)])
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


@pytest.mark.parametrize("L,limit", [(11, 1 << 43), (8, 1 << 44)])
def test_block_masks_equal_exact_divisibility(L, limit):
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
