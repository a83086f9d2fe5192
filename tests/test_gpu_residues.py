"""GPU residue parity: the value each lane carries after propagating over the
index range, S(n1-1) mod p, must equal the oracle's for every prime.

Zero residues (the factors) are rare for large primes, so factor-list
comparisons alone leave most of the arithmetic unobserved; the final residue
exposes every step.  Each case runs the default kernels (FP64 forms below
2^48), the integer forms only, and the generic (64-bit Montgomery) kernel.
"""
import random

import numpy as np
import pytest
import sympy

from cutrialdive_amd import _native as nat
from oracle import oracle as o

pytestmark = pytest.mark.gpu

FAMILIES = [
    (o.MERSENNE, 0),
    (o.PROTH, 3),
    (o.PROTH, 2718281828459),
    (o.RIESEL, 37),
    (o.RIESEL, (1 << 64) - 59),
    (o.PROTH_FIXED_N, 100000),
    (o.RIESEL_FIXED_N, 777),
    (o.FIBONACCI, 0),
    (o.LUCAS, 0),
    (o.TRIBONACCI, 0),
    (o.SMARANDACHE, 10),
    (o.SMARANDACHE, 16),
    (o.SMARANDACHE, 2),
    (o.SMARANDACHE, 3),
    (o.SMARANDACHE, 7),
    (o.SMARANDACHE, 64),
    (o.PEANO, 0),
]

# prime-size classes: 32-bit residues, 2^31..2^32, the narrow Smarandache
# range (FP64 Smarandache from 2^35), the 12-index doubling stride (2^37..2^41),
# the config-3/4 range, both sides of the FP64 limit 2^48, and near 2^63 /
# 2^64 (generic only above 2^63)
WINDOWS = [(3, 1 << 20), (1 << 31, 1 << 32), (1 << 32, 1 << 36), (1 << 37, 1 << 41), (1 << 38, 1 << 43),
           (1 << 47, 1 << 48), (1 << 48, 1 << 49), (1 << 62, 1 << 63), ((1 << 63) + 1, (1 << 64) - 1)]


def random_primes(lo, hi, k, seed):
    rng = random.Random(seed)
    ps = set()
    while len(ps) < k:
        ps.add(int(sympy.randprime(lo, hi)) if hi - lo > 1000 else int(sympy.nextprime(rng.randrange(lo, hi))))
    return sorted(p for p in ps if p > 2)


@pytest.mark.parametrize("seq,param", FAMILIES)
@pytest.mark.parametrize("win", range(len(WINDOWS)))
def test_final_residues(engine, seq, param, win):
    lo, hi = WINDOWS[win]
    primes = random_primes(lo, hi, 256, seed=seq * 97 + win)
    n0, n1 = (5, 2100) if seq != o.PEANO else (10 ** 12, 10 ** 12 + 500)
    want = o.residues_last(seq, param, n0, n1 - n0, primes)
    flag_sets = [0, nat.CTD_RUN_NO_FP64, nat.CTD_RUN_GENERIC_ONLY]
    if seq in (o.PROTH_FIXED_N, o.RIESEL_FIXED_N):  # the stepped stride loops as well
        flag_sets += [nat.CTD_RUN_STEPPED, nat.CTD_RUN_STEPPED | nat.CTD_RUN_NO_FP64]
    for flags in flag_sets:
        got = engine.residues(seq, param, n0, n1, primes, flags)
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, (flags, [(primes[i], int(got[i]), int(want[i])) for i in bad[:5]])


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_final_residues_jump_ahead(engine, seq, param):
    # start far from 1: the jump-ahead initial state must equal propagation
    primes = random_primes(1 << 40, 1 << 41, 128, seed=seq + 5)
    n0, n1 = 98765, 100321
    want = o.residues_last(seq, param, n0, n1 - n0, primes)
    got = engine.residues(seq, param, n0, n1, primes, 0)
    assert np.array_equal(got, want)


def test_smarandache_base10_digit_boundaries(engine):
    # every digit-length change from 1 to 6 digits inside one run
    primes = random_primes(1 << 32, 1 << 34, 256, seed=1)
    for n0, n1 in ((1, 120), (95, 1005), (9990, 10011), (99990, 100012), (1, 20000)):
        want = o.residues_last(o.SMARANDACHE, 10, n0, n1 - n0, primes)
        for flags in (0, nat.CTD_RUN_SM_NO_PAIR):
            got = engine.residues(o.SMARANDACHE, 10, n0, n1, primes, flags)
            assert np.array_equal(got, want), (n0, n1, flags)


def test_residues_rejects_even_moduli(engine):
    with pytest.raises(RuntimeError):
        engine.residues(o.MERSENNE, 0, 1, 10, [4], 0)


def _digits_multiplier(base, n):
    m = base
    while m <= n:
        m *= base
    return m


@pytest.mark.parametrize("base", [10, 2, 16, 7])
def test_smarandache_short_step_limit(engine, base):
    # The short Smarandache step needs p (0.75 m + 1) < 2^53 for the digit
    # multiplier m of the segment; a wave takes it only when all its lanes
    # qualify.  For each multiplier of the run: a wave of primes just below
    # the limit, a wave just above it, and waves that interleave primes below
    # the limit with primes at 2-4 times it (where the short step would be
    # inexact).  Final residues are compared by default (the short step below
    # its limit, the shifted step above it), with the shifted step wherever it
    # is exact (CTD_RUN_SM_NO_SHORT), without it (CTD_RUN_SM_NO_SHIFT: the
    # exact-product step above the limit) and with the general step only.
    n0, n1 = 5, 2100
    mults = sorted({_digits_multiplier(base, n) for n in range(n0, n1)})
    fp_max = (1 << 48) - 1
    rng = random.Random(base)
    for m in mults:
        lim = int((2 ** 53 - 2 ** 30) / (0.75 * m + 1))
        if lim >= fp_max or lim < 16 * n1:
            continue
        below, above, far = [], [], []
        p = lim
        while len(below) < 96:
            p = int(sympy.prevprime(p))
            below.append(p)
        p = lim
        while len(above) < 64 and p < fp_max:
            p = int(sympy.nextprime(p))
            if p < fp_max:
                above.append(p)
        far_hi = min(4 * lim, fp_max - 1000)
        while 2 * lim < far_hi and len(far) < 32:
            q = int(sympy.nextprime(rng.randrange(2 * lim, far_hi)))
            if q < fp_max:
                far.append(q)
        mixed = [x for pair in zip(below[64:], far) for x in pair] if far else below[64:]
        primes = below[:64] + above + mixed
        want = o.residues_last(o.SMARANDACHE, base, n0, n1 - n0, primes)
        _check_sm_forms(engine, base, n0, n1, primes, want, m,
                        (0, nat.CTD_RUN_SM_NO_SHORT, nat.CTD_RUN_SM_NO_SHIFT, nat.CTD_RUN_SM_FULL_STEP))


def _check_sm_forms(engine, base, n0, n1, primes, want, tag, flag_sets):
    """Final residues of every 64-prime wave of `primes` in each step form,
    one prime per lane (CTD_RUN_SM_NO_PAIR) and two: the two-prime kernel
    pairs the waves 2i and 2i+1 of its input, so there every wave is given
    twice in a row and a pair holds the same primes (its step choices then
    follow that wave alone)."""
    waves = [primes[i:i + 64] for i in range(0, len(primes), 64)]
    doubled = [x for w in waves for x in w + w]
    ref = dict(zip(primes, (int(x) for x in want)))
    want2 = np.array([ref[x] for x in doubled], dtype=np.uint64)
    for flags in flag_sets:
        for ps, exp, fl in ((primes, want, flags | nat.CTD_RUN_SM_NO_PAIR), (doubled, want2, flags)):
            got = engine.residues(o.SMARANDACHE, base, n0, n1, ps, fl)
            bad = np.nonzero(got != exp)[0]
            assert len(bad) == 0, (tag, fl, [(ps[i], int(got[i]), int(exp[i])) for i in bad[:5]])


def _odd_part(m):
    while m % 2 == 0:
        m //= 2
    return m


@pytest.mark.parametrize("base", [10, 12, 6, 14, 2, 3])
def test_smarandache_exact_product_limit(engine, base):
    # The exact-product Smarandache step needs (p / 2 + n1) o < 2^53 for the
    # odd part o of the digit multiplier m (every state is at most
    # (1/2 + 2^-30) p + n1 in size, so r m is exact); a wave takes it when all
    # its lanes qualify and the short step does not.  For each
    # multiplier whose limit is below 2^48: a wave of primes just below the
    # limit, a wave just above it, and waves that interleave primes below it
    # with primes at 2-4 times it.  Power-of-two bases (o = 1) qualify at
    # every prime below 2^48: waves spread over [2^40, 2^48).  The shifted
    # step has the same limit.  Final residues with every step allowed, with
    # the shifted step wherever it is exact (CTD_RUN_SM_NO_SHORT), without it
    # (CTD_RUN_SM_NO_SHIFT: the short step below its own limit, else the
    # exact-product step), without the shifted and the short step, and with
    # the general step only.
    n0, n1 = 5, 2100
    mults = sorted({_digits_multiplier(base, n) for n in range(n0, n1)})
    fp_max = (1 << 48) - 1
    rng = random.Random(base + 100)
    sets = []
    for m in mults:
        lim = int(((2 ** 53 - 2 ** 30) / _odd_part(m) - n1) / 0.5000001)
        if lim >= fp_max:
            continue
        below, above, far = [], [], []
        p = lim
        while len(below) < 96:
            p = int(sympy.prevprime(p))
            below.append(p)
        p = lim
        while len(above) < 64 and p < fp_max:
            p = int(sympy.nextprime(p))
            if p < fp_max:
                above.append(p)
        far_hi = min(4 * lim, fp_max - 1000)
        while 2 * lim < far_hi and len(far) < 32:
            q = int(sympy.nextprime(rng.randrange(2 * lim, far_hi)))
            if q < fp_max:
                far.append(q)
        mixed = [x for pair in zip(below[64:], far) for x in pair] if far else below[64:]
        sets.append((m, below[:64] + above + mixed))
    if _odd_part(base) == 1:
        spread = sorted(int(sympy.nextprime(rng.randrange(1 << 40, fp_max - 1000))) for _ in range(192))
        sets.append((0, spread))
    assert sets
    for m, primes in sets:
        want = o.residues_last(o.SMARANDACHE, base, n0, n1 - n0, primes)
        _check_sm_forms(engine, base, n0, n1, primes, want, m,
                        (0, nat.CTD_RUN_SM_NO_SHORT, nat.CTD_RUN_SM_NO_SHIFT,
                         nat.CTD_RUN_SM_NO_SHIFT | nat.CTD_RUN_SM_NO_SHORT, nat.CTD_RUN_SM_FULL_STEP))


@pytest.mark.parametrize("lim", [1 << 43, 1 << 44, 1 << 45, 1 << 46, 1 << 47, 1 << 48])
def test_tribonacci_lazy_form_limit(engine, lim):
    # The lazy Tribonacci form takes blocks of 11 indices on waves whose
    # primes are all below 2^43, then 10, 8, 7, 6 and 5 below 2^44 .. 2^48
    # (the integer form above); a wave of primes just below each limit, one
    # just above, and waves that interleave both (which take the next form).
    below, above = [], []
    p = lim
    while len(below) < 96:
        p = int(sympy.prevprime(p))
        below.append(p)
    p = lim
    while len(above) < 96:
        p = int(sympy.nextprime(p))
        above.append(p)
    mixed = [x for pair in zip(below[64:], above[64:]) for x in pair]
    primes = below[:64] + above[:64] + mixed
    for n0, n1 in ((5, 2100), (1, 9), (3, 11), (1000, 1017), (7, 18), (2, 27)):
        want = o.residues_last(o.TRIBONACCI, 0, n0, n1 - n0, primes)
        got = engine.residues(o.TRIBONACCI, 0, n0, n1, primes, 0)
        assert np.array_equal(got, want), (lim, n0, n1)
