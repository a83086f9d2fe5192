"""Pins the oracle (oracle/ctd_oracle.c) before the GPU results are checked against it.

Every function is checked against (a) the reference's own known-answer
vectors (tests/golden/reference_kats.json, transcribed from the reference's
tests and README) and (b) an independent computation: Python big integers
for the sequence values, sympy's primality for the sieve, pow() for powers.
"""
import json
import os
import random

import numpy as np
import pytest
import sympy

from oracle import oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))


def ndigits(n, b):
    """Number of base-b digits of n >= 1, by repeated division."""
    d = 0
    while n:
        n //= b
        d += 1
    return d


def u128(v):
    v = int(v)
    return [v & ((1 << 64) - 1), v >> 64]


def arr(xs):
    return (o.u64 * len(xs))(*xs)


# ---------------------------------------------------------------- arithmetic


def test_mu64_kat():
    for p, mu in KATS["mu64"]:
        assert o.lib().ctdo_mu64(p) == mu, p


def test_mu128_kat():
    for p, mu in KATS["mu128"]:
        out = (o.u64 * 2)()
        o.lib().ctdo_mu128(p, out)
        assert out[0] + (out[1] << 64) == int(mu), p


def test_mul128hi_kat_and_bruteforce():
    L = o.lib()
    rows = [tuple(int(x) for x in r) for r in KATS["mul128hi"]]
    rng = random.Random(7)
    rows += [(a, b, (a * b) >> 128) for a, b in ((rng.getrandbits(128), rng.getrandbits(128)) for _ in range(500))]
    for a, b, hi in rows:
        out = (o.u64 * 2)()
        L.ctdo_mul128hi(arr(u128(a)), arr(u128(b)), out)
        assert out[0] + (out[1] << 64) == hi


def test_reciprocal_mod2by1_modnby1_kat():
    L = o.lib()
    for d, r in KATS["reciprocal"]:
        assert L.ctdo_reciprocal(d) == r
        assert r == ((1 << 128) - 1) // d - (1 << 64)  # the definition
    for u1, u0, d, rnd, exp in KATS["mod2by1"]:
        assert L.ctdo_mod2by1(u1, u0, d, rnd) == exp
        assert exp == ((u1 << 64) | u0) % d
    nums = KATS["modnby1"]
    for name, d, exp in nums["cases"]:
        limbs = nums[name]
        assert L.ctdo_modnby1(arr(limbs), len(limbs), d) == exp
        v = sum(x << (64 * i) for i, x in enumerate(limbs))
        assert v % d == exp


def test_mulmod_modpow_bruteforce():
    L = o.lib()
    rng = random.Random(11)
    for _ in range(3000):
        d = rng.getrandbits(rng.randint(2, 64)) | 1
        if d < 3:
            continue
        a, b = rng.randrange(d), rng.getrandbits(64)  # b >= d occurs (mulmod(r, k) in riesel_proth.hpp:95)
        assert L.ctdo_mulmod(a, b, d) == (a * b) % d
        a2, b2 = rng.randrange(d), rng.randrange(d)
        assert L.ctdo_mulmod(a2, b2, d) == (a2 * b2) % d
        e = rng.getrandbits(rng.randint(1, 64))
        assert L.ctdo_modpow(a2, e, d) == pow(a2, e, d)
        mu = L.ctdo_mu64(d)
        x = rng.getrandbits(64)
        assert L.ctdo_mod(x, d, mu) == x % d


def test_digits_in_base():
    L = o.lib()
    for b in range(2, 65):
        for n in list(range(1, 300)) + [b ** k + j for k in range(1, 8) for j in (-1, 0, 1) if b ** k < 2 ** 63]:
            want = ndigits(n, b)
            assert L.ctdo_digits_in_base(b, n) == want, (b, n)


# ---------------------------------------------------------------- the sieve


@pytest.mark.parametrize("w", KATS["sieve_windows"], ids=lambda w: f"{w['start']}")
def test_sieve_autotest_windows(w):
    got = o.sieve(w["start"], w["end"]).tolist()
    if "primes" in w:
        assert got == w["primes"]
    if "count" in w:
        assert len(got) == w["count"]


def test_sieve_bruteforce_windows():
    rng = random.Random(3)
    for _ in range(40):
        a = rng.getrandbits(rng.randint(1, 40))
        b = a + rng.randint(0, 5000)
        got = o.sieve(a, b).tolist()
        want = list(sympy.primerange(a, b))
        assert got == want, (a, b)


def test_sieve_top_of_64_bit_range():
    # the ten largest primes below 2^64 are 2^64 - d for d in 59, 83, 95, 179,
    # 189, 257, 279, 323, 353, 363 (the Prime Pages' table of primes just less
    # than a power of two); --tf-bits 64 sieves up to f1 = 2^64 - 1
    top = [(1 << 64) - d for d in (363, 353, 323, 279, 257, 189, 179, 95, 83, 59)]
    assert o.sieve((1 << 64) - 400, (1 << 64) - 1).tolist() == top
    a = (1 << 64) - 5000
    assert o.sieve(a, (1 << 64) - 1).tolist() == list(sympy.primerange(a, (1 << 64) - 1))
    # 2^64 - 873 = 361373 x 51046... is odd with no prime factor below 8192
    c = (1 << 64) - 873
    assert min(sympy.factorint(c)) == 361373 and c not in o.sieve(c - 10, c + 10).tolist()


def test_trial_factor_near_2_64_planted_hit():
    # k = -2^-n mod p makes p divide k 2^n + 1 (Proth); the oracle must list p
    # at n and, checked by brute force, nothing else in the window
    p = (1 << 64) - 59
    n = 1000
    k = (-pow(2, -n, p)) % p
    f0, f1 = (1 << 64) - 4000, (1 << 64) - 1
    got = o.trial_factor(o.PROTH, k, n - 20, n + 20, f0, f1, 8)
    primes = list(sympy.primerange(f0, f1))
    want = [[q for q in primes if (k * pow(2, m, q) + 1) % q == 0] for m in range(n - 20, n + 20)]
    assert got == want and got[20] == [p]


def test_mersenne_orders_tool_equals_the_oracle(mersenne_lists):
    # the order-based Mersenne lists (the full-size config-1 check of the GPU
    # tests) against the oracle's trial factoring on a smaller case
    want = o.trial_factor(o.MERSENNE, 0, 2, 2000, 0, 1 << 20)
    assert mersenne_lists(20, 2000) == want
    assert mersenne_lists(20, 2000, n0=3)[0] == [7]  # M_3 = 7 lists itself
    assert mersenne_lists(20, 2000)[11 - 2] == [23, 89]


def test_prime_count_2_30_and_readme_primes():
    assert o.prime_count(0, 1 << 30) == KATS["prime_counts"]["30"]
    assert int(o.sieve((1 << 37) - 200, 1 << 37)[-1]) == KATS["readme_largest_prime_below_2_37"]
    assert int(o.sieve((1 << 38) - 200, 1 << 38)[-1]) == KATS["readme_largest_prime_below_2_38"]


# ---------------------------------------------------------------- families


def exact_value(seq, param, n):
    """S(n) as a Python integer, from each family's definition."""
    if seq == o.MERSENNE:
        return 2 ** n - 1
    if seq == o.PROTH:
        return param * 2 ** n + 1
    if seq == o.RIESEL:
        return param * 2 ** n - 1
    if seq == o.PROTH_FIXED_N:
        return n * 2 ** param + 1
    if seq == o.RIESEL_FIXED_N:
        return n * 2 ** param - 1
    if seq == o.SMARANDACHE:
        v = 0
        for i in range(1, n + 1):
            v = v * param ** ndigits(i, param) + i
        return v
    if seq in (o.FIBONACCI, o.LUCAS):
        a, b = (0, 1) if seq == o.FIBONACCI else (2, 1)
        for _ in range(n):
            a, b = b, a + b
        return a
    if seq == o.TRIBONACCI:
        a, b, c = 0, 0, 1
        for _ in range(n):
            a, b, c = b, c, a + b + c
        return a
    if seq == o.PEANO:
        return n
    raise ValueError(seq)


FAMS = [(o.MERSENNE, 0), (o.PROTH, 3), (o.PROTH, 1279), (o.RIESEL, 37), (o.RIESEL, 1), (o.PROTH_FIXED_N, 31),
        (o.RIESEL_FIXED_N, 100), (o.SMARANDACHE, 10), (o.SMARANDACHE, 2), (o.SMARANDACHE, 3), (o.SMARANDACHE, 16),
        (o.SMARANDACHE, 64), (o.FIBONACCI, 0), (o.LUCAS, 0), (o.TRIBONACCI, 0), (o.PEANO, 0)]


@pytest.mark.parametrize("seq,param", FAMS)
def test_residues_bruteforce(seq, param):
    rng = random.Random(seq * 1000 + param)
    for _ in range(6):
        d = int(sympy.randprime(3, 1 << rng.choice([8, 20, 33, 50, 63, 64])))
        n0 = rng.randint(1, 120)
        cnt = 60
        got = o.residues(seq, param, n0, cnt, d)
        want = [exact_value(seq, param, n) % d for n in range(n0, n0 + cnt)]
        assert got == want, (seq, param, d, n0)


def test_fibonacci_lucas_n0_zero_quirk():
    # fibonacci.hpp:68-71 starts n = 0 from {previous, current} = {0, 0}
    # (lucas.hpp:54-66 maps it to {0, 0} too): every later residue is 0
    for seq in (o.FIBONACCI, o.LUCAS):
        assert o.state(seq, 0, 0, 101) == [0, 0]
        assert o.residues(seq, 0, 0, 30, 101) == [0] * 30


def test_value_mod_2_propagation_equals_closed_form():
    # the GPU evaluates value_mod_2 per index; the reference propagates it
    L = o.lib()
    for seq, param in FAMS + [(o.PROTH, 2), (o.RIESEL, 2), (o.SMARANDACHE, 7)]:
        r = L.ctdo_value_mod_2(seq, param, 0)
        for n in range(0, 10000):
            assert r == L.ctdo_value_mod_2(seq, param, n), (seq, param, n)
            if n <= 200:
                assert r == exact_value(seq, param, n) % 2, (seq, param, n)
            r = L.ctdo_next_value_mod_2(seq, param, r, n)


def test_family_kats():
    K = KATS
    for n, d, exp in K["mersenne_value_mod"]:
        assert o.residues(o.MERSENNE, 0, n, 1, d)[0] == exp, (n, d)
    for n, exp in K["mersenne_value_mod_2"]:
        assert o.lib().ctdo_value_mod_2(o.MERSENNE, 0, n) == exp
    for k, n, d, exp in K["proth_value_mod"]:
        assert o.residues(o.PROTH, k, n, 1, d)[0] == exp
    for k, n, exp in K["proth_value_mod_2"]:
        assert o.lib().ctdo_value_mod_2(o.PROTH, k, n) == exp
    pn = K["proth_next"]
    assert o.residues(o.PROTH, pn["k"], pn["n0"], 9, pn["d"])[1:] == pn["seq"]
    for k, n, d, exp in K["riesel_value_mod"]:
        assert o.residues(o.RIESEL, k, n, 1, d)[0] == exp
    rn = K["riesel_next"]
    assert o.residues(o.RIESEL, rn["k"], rn["n0"], 9, rn["d"])[1:] == rn["seq"]
    for n, d, prev, cur in K["fibonacci_state"]:
        assert o.state(o.FIBONACCI, 0, n, d) == [prev, cur]
    fw = K["fibonacci_walk"]
    assert o.residues(o.FIBONACCI, 0, fw["n0"], len(fw["seq"]), fw["d"]) == fw["seq"]
    for n, d, prev, cur in K["lucas_state"]:
        assert o.state(o.LUCAS, 0, n, d) == [prev, cur]
    lw = K["lucas_walk"]
    assert o.residues(o.LUCAS, 0, lw["n0"], len(lw["seq"]), lw["d"]) == lw["seq"]
    for n, d, a, b, c in K["tribonacci_state"]:
        assert o.state(o.TRIBONACCI, 0, n, d) == [a, b, c]
    for n, v in K["tribonacci_value_mod_2"].items():
        assert o.lib().ctdo_value_mod_2(o.TRIBONACCI, 0, int(n)) == v
    for name, seq in (("fibonacci_values", o.FIBONACCI), ("lucas_values", o.LUCAS), ("tribonacci_values", o.TRIBONACCI)):
        for n, v in K[name].items():
            assert exact_value(seq, 0, int(n)) == v  # the brute force agrees with the reference's values
            d = 1000000007
            if int(n) > 0:
                assert o.residues(seq, 0, int(n), 1, d)[0] == v % d
    for n, v in K["smarandache_values_base10"].items():
        assert o.smarandache_value(10, int(n)) == v
    for n, v in K["smarandache_mod_2_base10"].items():
        assert o.lib().ctdo_value_mod_2(o.SMARANDACHE, 10, int(n)) == v
    for n, v in K["smarandache_mod_2_base3"].items():
        assert o.lib().ctdo_value_mod_2(o.SMARANDACHE, 3, int(n)) == v
    for b, r, n, d, exp in K["smarandache_next_mu"]:
        assert o.next_residue(o.SMARANDACHE, b, r, n, d) == exp, (b, r, n, d)


@pytest.mark.parametrize("base", [2, 3, 7, 10, 16, 19, 36, 64])
def test_smarandache_value_concatenation(base):
    # value(n) equals the naive concatenation (smarandache_tests.cpp:22-35)
    for n in list(range(0, 40)) + [base ** 2 - 1, base ** 2, 300]:
        assert o.smarandache_value(base, n) == exact_value(o.SMARANDACHE, base, n), (base, n)


# ---------------------------------------------------------------- end to end


def test_tf_end_to_end_kats():
    for case in KATS["tf_end_to_end"]:
        got = o.trial_factor(case["seq"], case["param"], case["n0"], case["n1"], case["f0"], case["f1"], case["max"])
        for n, fs in case["expect"].items():
            n = int(n)
            idx = n - case["n0"] if n >= case["n0"] else n
            assert got[idx] == fs, (case["what"], n)


def test_readme_smarandache_lines():
    # README.md:64-70 lists every factor below 2^37; all of them are below 2^24
    want = {int(k): v for k, v in KATS["readme_sm10_to_2_37"].items()}
    got = o.trial_factor(o.SMARANDACHE, 10, 10, 13, 0, 1 << 32)
    for n in (10, 11, 12):
        assert got[n - 10] == want[n]
    got = o.trial_factor(o.SMARANDACHE, 10, 997, 1000, 0, 1 << 24)
    for n in (997, 998, 999):
        assert got[n - 997] == want[n]
        v = exact_value(o.SMARANDACHE, 10, n)
        assert all(v % p == 0 for p in want[n])


def test_tf_matches_bruteforce_small():
    # every index, every prime: the factor lists equal trial division of the exact values
    for seq, param in FAMS:
        n0, n1 = (1, 40) if seq not in (o.PROTH_FIXED_N, o.RIESEL_FIXED_N) else (2, 40)
        if seq == o.TRIBONACCI:
            n0 = 2
        f1 = 3000
        got = o.trial_factor(seq, param, n0, n1, 0, f1, 4096)
        primes = list(sympy.primerange(2, f1))
        for i, n in enumerate(range(n0, n1)):
            v = exact_value(seq, param, n)
            want = [p for p in primes if v % p == 0]
            assert got[i] == want, (seq, param, n)


@pytest.mark.parametrize("f0,f1", [(0, 2), (1, 2), (0, 1), (2, 3), (0, 3), (1, 8), (3, 4)])
def test_tf_tiny_prime_windows_match_bruteforce(f0, f1):
    # windows that hold no prime, only 2, only 3: no other number is ever listed
    for seq, param in FAMS:
        n0, n1 = (2, 30)
        got = o.trial_factor(seq, param, n0, n1, f0, f1, 64)
        primes = list(sympy.primerange(f0, f1))
        for i, n in enumerate(range(n0, n1)):
            v = exact_value(seq, param, n)
            assert got[i] == [p for p in primes if v % p == 0], (seq, param, n, f0, f1)


def test_tf_threads_equivalent():
    a = o.trial_factor(o.PROTH, 1279, 0, 300, 0, 1 << 20, threads=1)
    b = o.trial_factor(o.PROTH, 1279, 0, 300, 0, 1 << 20, threads=8)
    assert a == b


def test_fib_large_factor_fixture():
    """tests/golden/fib_large_factors.json: each q divides F(m) (Python
    integers), and the oracle lists q at index m."""
    import json

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fib_large_factors.json")) as f:
        pairs = json.load(f)["pairs"]
    assert len(pairs) >= 10
    F = [0, 1]
    for _ in range(2, 1500):
        F.append(F[-1] + F[-2])
    for m, q in pairs[:6]:
        assert F[m] % q == 0
        r = o.trial_factor(o.FIBONACCI, 0, m - 2, m + 3, q - 100, q + 100)
        assert q in r[2]
