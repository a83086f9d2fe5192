"""GPU parity: the HIP engine (through the C ABI) against the oracle.

Factor lists must be bit-identical: same primes per index, same counts.
Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import cutrialdive_amd as ctd
from cutrialdive_amd import _native as nat
from oracle import oracle as o

pytestmark = pytest.mark.gpu

FAMILIES = [
    (o.MERSENNE, 0),
    (o.PROTH, 1279),
    (o.RIESEL, 37),
    (o.PROTH, 3),
    (o.RIESEL, 78557),
    (o.SMARANDACHE, 10),
    (o.SMARANDACHE, 16),
    (o.SMARANDACHE, 3),
    (o.SMARANDACHE, 2),
    (o.SMARANDACHE, 64),
    (o.FIBONACCI, 0),
    (o.LUCAS, 0),
    (o.PROTH_FIXED_N, 1000),
    (o.RIESEL_FIXED_N, 333),
    (o.TRIBONACCI, 0),
    (o.PEANO, 0),
]


def gpu_lists(eng, seq, param, n0, n1, f0, f1, mx=1024, **kw):
    opts = None
    if kw.get("flags") is None:
        kw.pop("flags", None)
    if kw:
        opts = nat.CtdRunOpts(kw.get("batch_length", 0), 0, 0, kw.get("chunk_length", 0), kw.get("flags", 0))
    counts, factors, last = eng.run_job(seq, param, n0, n1, f0, f1, mx, opts=opts)
    res = ctd.FactoringResults.from_buffer(n0, counts, factors, mx)
    return res.factors, counts, last


def oracle_counts(seq, param, n0, n1, f0, f1, mx=1024):
    c, _ = o.trial_factor_raw(seq, param, n0, n1, f0, f1, mx)
    return c


def assert_same(g, gc, seq, param, n0, n1, f0, f1, mx=1024):
    """Counts equal everywhere; stored lists equal wherever they did not
    overflow max_factors (beyond it the reference stores an arbitrary subset,
    factors_buffer.hpp:110-121)."""
    rc, rf = o.trial_factor_raw(seq, param, n0, n1, f0, f1, mx)
    assert np.array_equal(gc, rc)
    r = [sorted(int(x) for x in rf[i * mx : i * mx + min(int(c), mx)]) for i, c in enumerate(rc)]
    for i, c in enumerate(rc):
        if c <= mx:
            assert g[i] == r[i], (n0 + i, g[i], r[i])


@pytest.mark.parametrize(
    "start,end",
    [
        (256, 307),
        (1 << 16, (1 << 16) + 100),
        ((1 << 31) - 1, (1 << 31) + 100),
        (3299100100, 3299100100 + 256),
        (1, 1 << 17),
        ((1 << 32) - 256, (1 << 32) + 256),
        (65537 * 65537 - 200, 65537 * 65537 + 1),
        ((1 << 34) - 200, (1 << 34) + 200),
        ((1 << 41) - (1 << 37) + (1 << 32) - 1, (1 << 41) - (1 << 37) + (1 << 32) + 1000),
        (0, 3),
        (2, 3),
        (1 << 40, (1 << 40) + (1 << 22) + 12345),
    ],
)
def test_sieve_windows(engine, start, end):
    g = engine.sieve(start, end)
    r = o.sieve(start, end)
    assert np.array_equal(g, r)


def test_sieve_pi_2_30(engine):
    assert len(engine.sieve(0, 1 << 30)) == 54400028


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_small(engine, seq, param):
    n0, n1 = (1, 300) if seq != o.PROTH_FIXED_N and seq != o.RIESEL_FIXED_N else (3, 400)
    f0, f1 = 0, 1 << 20
    g, gc, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1)
    # tribonacci T(1) = 0: index 0 overflows max_factors
    assert_same(g, gc, seq, param, n0, n1, f0, f1)


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_generic_kernel(engine, seq, param):
    n0, n1 = 50, 250
    f0, f1 = 1 << 20, (1 << 20) + (1 << 18)
    g, _, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1, flags=nat.CTD_RUN_GENERIC_ONLY)
    r = o.trial_factor(seq, param, n0, n1, f0, f1)
    assert g == r


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_w64_window(engine, seq, param):
    # primes above 2^32: the 64-bit step paths
    n0, n1 = 100, 400
    f0, f1 = 1 << 36, (1 << 36) + (1 << 16)
    g, _, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1)
    r = o.trial_factor(seq, param, n0, n1, f0, f1)
    assert g == r


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_w64_window_integer_forms(engine, seq, param):
    # the same window without the FP64 forms
    n0, n1 = 100, 400
    f0, f1 = 1 << 36, (1 << 36) + (1 << 16)
    g, _, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1, flags=nat.CTD_RUN_NO_FP64)
    r = o.trial_factor(seq, param, n0, n1, f0, f1)
    assert g == r


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_w42_window(engine, seq, param):
    # primes near 2^42 with large indices: FP64 set-up and loops
    n0, n1 = 40000, 40600
    f0, f1 = (1 << 42) + (1 << 30), (1 << 42) + (1 << 30) + (1 << 17)
    g, _, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1)
    r = o.trial_factor(seq, param, n0, n1, f0, f1)
    assert g == r


@pytest.mark.parametrize("seq,param", FAMILIES)
def test_family_chunked(engine, seq, param):
    # index chunks start mid-range: jump-ahead must equal propagation
    n0, n1 = 7, 3000
    f0, f1 = 1000, 20000
    g, _, _ = gpu_lists(engine, seq, param, n0, n1, f0, f1, chunk_length=97)
    r = o.trial_factor(seq, param, n0, n1, f0, f1)
    assert g == r


@pytest.mark.parametrize("block", [64, 128, 192, 256])
def test_block_and_grid_size_flags(engine, block):
    # --block-size / --grid-size reach the residue launch as given
    # (trial_factoring.cuh:280-286); the kernels are grid-stride loops, so any
    # grid computes the same factors
    seq, param, n0, n1, f0, f1 = o.PROTH, 1279, 0, 600, 0, 1 << 20
    r = o.trial_factor(seq, param, n0, n1, f0, f1, 64)
    for grid in (0, 1, 7, 5000):
        opts = nat.CtdRunOpts(0, grid, block, 0, 0)
        counts, factors, _ = engine.run_job(seq, param, n0, n1, f0, f1, 64, opts=opts)
        assert ctd.FactoringResults.from_buffer(n0, counts, factors, 64).factors == r, (block, grid)


@pytest.mark.parametrize("block", [32, 100, 320, 512, 1024])
def test_unsupported_block_size_is_an_error(engine, block):
    # the reference hands any block size to the launch; here the kernels take
    # whole 64-lane waves up to 256 lanes, and other sizes fail before any work
    with pytest.raises(RuntimeError, match="block size"):
        engine.run_job(o.PROTH, 1279, 0, 80, 0, 1 << 16, 8, opts=nat.CtdRunOpts(0, 0, block, 0, 0))


# Primes whose order of 2 is below 64 repeat stride targets: p | 2^59 - 1
# (order 59) and p | 2^47 - 1 (order 47), both below 2^48.
SMALL_ORDER_PRIMES = [(3203431780337, 59), (13264529, 47)]


@pytest.mark.parametrize("p,order", SMALL_ORDER_PRIMES)
@pytest.mark.parametrize("flags", [None, nat.CTD_RUN_NO_FP64])
def test_doubling_small_order_primes(engine, p, order, flags):
    """Mersenne: p | 2^n - 1 exactly when order | n; Proth k = p - 1 (k 2^n + 1
    = 1 - 2^n mod p) likewise.  The 64 stride targets repeat (j and j + order
    give the same value) and every multiple of the order must be found, in
    64-index blocks and in the tail."""
    f0, f1 = p - 2000, p + 2000
    for seq, k in ((o.MERSENNE, 0), (o.PROTH, p - 1), (o.RIESEL, 1)):
        n0, n1 = 3, 3 + 64 * 37 + 29
        r = o.trial_factor(seq, k, n0, n1, f0, f1, 64)
        assert sum(1 for i, fs in enumerate(r) if p in fs) == len([n for n in range(n0, n1) if n % order == 0])
        g, _, _ = gpu_lists(engine, seq, k, n0, n1, f0, f1, mx=64, flags=flags)
        assert g == r, (seq, flags)


@pytest.mark.parametrize("seq", [o.PROTH_FIXED_N, o.RIESEL_FIXED_N])
@pytest.mark.parametrize("n,k0,N,f0,f1", [
    (1000, 0, 20000, 0, 1 << 16),                 # small primes: many hits per prime, k = 0 included
    (7, 1, 300000, 3, 2000),                      # more indices than the primes: a class per prime
    # the rest: a hit planted at offset 100 for a prime p0 of the window
    (100000, None, 5000, 1 << 40, (1 << 40) + (1 << 20)),
    (1, None, 777, (1 << 36) - 5000, (1 << 36) + 5000),
    (64, 1 << 53, 600, 1 << 30, (1 << 30) + (1 << 16)),                   # indices past 2^52
    (333, None, 4000, (1 << 48) - (1 << 16), (1 << 48) + (1 << 16)),      # both sides of the FP64 limit
    (10 ** 12, None, 3000, (1 << 47) + 12345, (1 << 47) + 12345 + (1 << 14)),
])
def test_fixed_n_closed_form(engine, seq, n, k0, N, f0, f1):
    """Fixed n: the hits of each prime are the k = -+2^-n (mod p) of the
    index range, solved once per prime; they equal the stepped forms' and the
    oracle's, also when a chunk length splits the index range."""
    planted = k0 is None or k0 >= 1 << 50
    if planted:
        primes = [int(x) for x in o.sieve(f0, f1)]
        p0 = primes[len(primes) // 2]
        inv = pow(2, -n, p0)
        ks = (p0 - inv) % p0 if seq == o.PROTH_FIXED_N else inv
        if k0:  # shift the solution up by multiples of p0 to just past k0
            ks += -(-(k0 - ks) // p0) * p0
        k0 = ks - 100
    k1 = k0 + N
    r = o.trial_factor(seq, n, k0, k1, f0, f1, 64, threads=16)
    if planted:
        assert p0 in r[100]
    for flags in (None, nat.CTD_RUN_STEPPED, nat.CTD_RUN_NO_FP64):
        g, _, _ = gpu_lists(engine, seq, n, k0, k1, f0, f1, mx=64, flags=flags)
        assert g == r, flags
    g, _, _ = gpu_lists(engine, seq, n, k0, k1, f0, f1, mx=64, chunk_length=97)
    assert g == r


def test_reference_tf_cases(engine):
    # tests/tf_tests.cpp:70-109
    g, _, _ = gpu_lists(engine, o.MERSENNE, 0, 1, 101, 0, 1 << 16)
    assert g[96] == [11447]
    g, _, _ = gpu_lists(engine, o.PROTH, 1279, 0, 80, 0, 1 << 16, 128)
    assert g[79] == [3, 2063]
    # tests/tf_tests.cpp:18-67 (S(n) = n)
    g, _, _ = gpu_lists(engine, o.PEANO, 0, 2956708207, 2956708707, 0, 1 << 24)
    assert g[0] == []
    assert g[1] == [2, 43, 4297541]
    assert g[2] == [3, 7, 13, 107, 127, 797]
    assert g[-1] == [2, 7, 4310071]


def test_readme_smarandache_lines(engine):
    # README.md:64-70, Sm10 to 2^37: checked on windows around each listed factor
    expect = {10: [2, 5, 1234567891], 11: [3, 7, 13, 67, 107, 630803], 12: [2, 3, 2437, 2110805449]}
    g, _, _ = gpu_lists(engine, o.SMARANDACHE, 10, 10, 13, 0, 1 << 32)
    for n, fs in expect.items():
        assert g[n - 10] == fs


def test_cfg1_reduced(engine):
    # Mersenne n in [2, 4000), p < 2^24 (config 1 at reduced size)
    g, gc, _ = gpu_lists(engine, o.MERSENNE, 0, 2, 4000, 0, 1 << 24)
    r = o.trial_factor(o.MERSENNE, 0, 2, 4000, 0, 1 << 24)
    assert g == r


def test_cfg1_full_size(engine, mersenne_lists):
    # config 1 at its full size, Mersenne n in [2, 20000) against every prime
    # below 2^30 (1.09e12 updates), checked index by index against the lists
    # the multiplicative orders of 2 give (p | M_n <=> ord_p(2) | n, found
    # without stepping any index, and pinned against the oracle on the CPU)
    print("[cfg1] mersenne n in [2, 20000), p < 2^30", flush=True)
    want = mersenne_lists(30, 20000)
    mx = 256
    assert max(len(x) for x in want) <= mx
    g, gc, last = gpu_lists(engine, o.MERSENNE, 0, 2, 20000, 0, 1 << 30, mx=mx)
    assert np.array_equal(gc, np.array([len(x) for x in want], dtype=gc.dtype))
    assert g == want
    assert last == (1 << 30) - 35  # the largest prime below 2^30


def test_cfg2_subwindow(engine):
    # Sm10 n in [10, 1000) over [2^32, 2^32 + 2^28) (config 2 sub-window,
    # 12.4 M primes); it holds zero residues, so the lists are not trivially empty
    f0, f1 = 1 << 32, (1 << 32) + (1 << 28)
    g, gc, _ = gpu_lists(engine, o.SMARANDACHE, 10, 10, 1000, f0, f1)
    r = o.trial_factor(o.SMARANDACHE, 10, 10, 1000, f0, f1, threads=16)
    assert g == r
    assert int(gc.sum()) >= 3


def test_cfg3_subwindow(engine):
    # Sm16 n in [1, 2000) and Proth k=3 n in [1, 20000) over [2^40, 2^40 + 2^20)
    f0, f1 = 1 << 40, (1 << 40) + (1 << 20)
    g, _, _ = gpu_lists(engine, o.SMARANDACHE, 16, 1, 2000, f0, f1)
    assert g == o.trial_factor(o.SMARANDACHE, 16, 1, 2000, f0, f1)
    g, _, _ = gpu_lists(engine, o.PROTH, 3, 1, 20000, f0, f1)
    assert g == o.trial_factor(o.PROTH, 3, 1, 20000, f0, f1)


def test_cfg3_proth_full_index_range(engine):
    # config 3b's whole index range, Proth k=3 n in [1, 10^5), on [2^40, 2^40 + 2^18)
    f0, f1 = 1 << 40, (1 << 40) + (1 << 18)
    print("[cfg3] proth,3 n in [1, 100000)", flush=True)
    g, _, _ = gpu_lists(engine, o.PROTH, 3, 1, 100000, f0, f1, mx=64)
    assert g == o.trial_factor(o.PROTH, 3, 1, 100000, f0, f1, 64, threads=16)


def test_cfg3_batched_beside_plain(engine):
    # config 3's two jobs in the batched form (several tiles each) and in the
    # plain form: both equal the oracle
    specs = [(o.SMARANDACHE, 16, 1, 2000, 1 << 40, (1 << 40) + (1 << 22)),
             (o.PROTH, 3, 1, 100000, 1 << 40, (1 << 40) + (1 << 18))]
    mx = 64
    jobs = [nat.CtdJob(s, mx, p, a, b, f0, f1) for s, p, a, b, f0, f1 in specs]
    print("[cfg3] batched sm16 + proth3", flush=True)
    res, st = engine.run_batch(jobs, tile_length=1 << 20, streams=2)
    assert st.tiles == 4 + 1
    for (s, p, a, b, f0, f1), (c, f) in zip(specs, res):
        r = o.trial_factor(s, p, a, b, f0, f1, mx, threads=16)
        assert ctd.FactoringResults.from_buffer(a, c, f, mx).factors == r, (s, p)
        g, _, _ = gpu_lists(engine, s, p, a, b, f0, f1, mx=mx)
        assert g == r, (s, p)


def test_n0_zero_quirks(engine):
    # Fibonacci / Lucas starting at n0 = 0 (fibonacci.hpp:68-71): every prime
    # divides every index, so the counts overflow max_factors identically
    for seq in (o.FIBONACCI, o.LUCAS, o.TRIBONACCI, o.MERSENNE, o.SMARANDACHE):
        param = 10 if seq == o.SMARANDACHE else 0
        _, gc, _ = gpu_lists(engine, seq, param, 0, 40, 0, 5000, mx=8)
        rc = oracle_counts(seq, param, 0, 40, 0, 5000, mx=8)
        assert np.array_equal(gc, rc), seq


def test_small_batches_and_resume_accumulate(engine):
    # several prime batches and a two-part run equal one run
    seq, param, n0, n1 = o.PROTH, 1279, 0, 500
    mx = 256
    full, fc, last = gpu_lists(engine, seq, param, n0, n1, 0, 1 << 22, mx=mx, batch_length=1 << 20)
    c1, f1_, last1 = engine.run_job(seq, param, n0, n1, 0, 1 << 21, mx)
    c2, f2_, last2 = engine.run_job(seq, param, n0, n1, last1 + 1, 1 << 22, mx, counts=c1, factors=f1_)
    res = ctd.FactoringResults.from_buffer(n0, c2, f2_, mx)
    assert res.factors == full
    assert last2 == last
    assert last1 == int(o.sieve((1 << 21) - 100, 1 << 21)[-1])


# Windows for the stride forms of the doubling and fixed-n families, from
# 2^36 to 2^48, including windows just below 2^37, 2^41 and 2^45 (earlier
# kernels changed form there; every window takes the word strides now).
STRIDE_WINDOWS = [(1 << 36) + 24690, (1 << 37) - 2000, (1 << 40) + (1 << 39), (1 << 41) - 2000,
                  (1 << 44) + (1 << 40), (1 << 45) - 2000, (1 << 47) + (1 << 30)]
# Offsets of the planted hit from n0: the check before the loop, the first,
# last and middle positions of the 128-index blocks of the doubling word
# stride (offsets 1..512 with N = 600), of its 64- and 16-index tail blocks
# (513..576, 577..592), of the 64- and 16-index blocks of the fixed-n stride,
# and the single-step tails.
STRIDE_OFFSETS = [0, 1, 2, 5, 11, 12, 13, 15, 16, 17, 63, 64, 65, 127, 128, 129, 255, 256, 257, 300, 383, 384,
                  511, 512, 513, 575, 576, 577, 588, 589, 591, 592, 593, 599]


@pytest.mark.parametrize("seq,flags", [(o.PROTH, None), (o.RIESEL, None), (o.PROTH_FIXED_N, None),
                                       (o.RIESEL_FIXED_N, None), (o.PROTH_FIXED_N, nat.CTD_RUN_STEPPED),
                                       (o.RIESEL_FIXED_N, nat.CTD_RUN_STEPPED)])
@pytest.mark.parametrize("w", STRIDE_WINDOWS)
def test_stride_planted_hits(engine, seq, flags, w):
    """At primes this large a random window holds no hits, so each job is
    built around one: p0 divides S at a chosen index.  Proth/Riesel:
    k = -+2^-n (mod p0) makes p0 | k 2^n +- 1; fixed n: the index k* = -+2^-n
    (mod p0) does."""
    f0, f1 = w, w + 4096
    primes = [int(x) for x in o.sieve(f0, f1)]
    p0 = primes[len(primes) // 3]
    N = 600
    for off in STRIDE_OFFSETS:
        if seq in (o.PROTH, o.RIESEL):
            n_hit = 1000 + 7 * off
            inv = pow(2, -n_hit, p0)
            param = (p0 - inv) % p0 if seq == o.PROTH else inv
            n0 = n_hit - off
        else:
            param = 777
            inv = pow(2, -param, p0)
            k_hit = (p0 - inv) % p0 if seq == o.PROTH_FIXED_N else inv
            k_hit += p0 * (1 + off % 3)  # index values above p0 as well
            n0 = k_hit - off
        g, gc, _ = gpu_lists(engine, seq, param, n0, n0 + N, f0, f1, flags=flags)
        assert p0 in g[off], (seq, w, off)
        assert_same(g, gc, seq, param, n0, n0 + N, f0, f1)


def _fib_factor_pairs():
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fib_large_factors.json")
    with open(path) as f:
        return [tuple(x) for x in json.load(f)["pairs"]]


@pytest.mark.parametrize("m,q", _fib_factor_pairs())
def test_theta_form_fibonacci_lucas_large_primes(engine, m, q):
    """The theta stride form (work items of at least 1024 indices) at primes
    2^37..2^47, around known factors q | F(m): q divides F at every multiple
    of its rank of apparition, and L at the odd multiples of half of it."""
    f0, f1 = q - 3000, q + 3000
    n0 = max(1, m - 600)
    g, gc, _ = gpu_lists(engine, o.FIBONACCI, 0, n0, n0 + 2000, f0, f1)
    assert q in g[m - n0]
    assert_same(g, gc, o.FIBONACCI, 0, n0, n0 + 2000, f0, f1)
    n0 = max(1, m // 2 - 500)
    g, gc, _ = gpu_lists(engine, o.LUCAS, 0, n0, n0 + 2000, f0, f1)
    assert_same(g, gc, o.LUCAS, 0, n0, n0 + 2000, f0, f1)


def _trib_factor_pairs():
    """One pair (n, q), q | T(n), per bit length of q from 32 to 48, and every
    pair with q in [2^42, 2^44), next to the limits of the lazy form's 11-
    and 8-index blocks."""
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trib_large_factors.json")
    with open(path) as f:
        pairs = [tuple(x) for x in json.load(f)["pairs"]]
    first = {}
    for n, q in pairs:
        first.setdefault(q.bit_length(), (n, q))
    return sorted(set(first.values()) | {(n, q) for n, q in pairs if 43 <= q.bit_length() <= 44},
                  key=lambda z: z[1])


@pytest.mark.parametrize("n,q", _trib_factor_pairs())
def test_tribonacci_planted_hits(engine, n, q):
    """q | T(n) planted at every position of a 30-index job: the check before
    the loop, each position of the lazy form's 11-index (q < 2^43) or 8-index
    (q < 2^44) blocks, or of the pair form's blocks (q >= 2^44), and the
    single-step tail."""
    T = [0, 0, 1]
    while len(T) <= n:
        T.append(T[-1] + T[-2] + T[-3])
    assert T[n] % q == 0
    f0, f1 = q - 600, q + 600
    N = 30
    lo = n - (N - 1)
    want = o.trial_factor(o.TRIBONACCI, 0, lo, n + N, f0, f1)
    for off in range(N):
        n0 = n - off
        g, _, _ = gpu_lists(engine, o.TRIBONACCI, 0, n0, n0 + N, f0, f1)
        assert q in g[off], (q, off)
        assert g == want[n0 - lo : n0 - lo + N], (q, off)


def _trib_near_triples():
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trib_near_residues.json")
    with open(path) as f:
        return [tuple(x) for x in json.load(f)["triples"]]


@pytest.mark.parametrize("n,r,q", _trib_near_triples())
def test_tribonacci_near_misses(engine, n, r, q):
    """T(n) = r (mod q) for a small r != 0 at q > |r| 2^34: wherever the lazy
    form leaves T(n) unreduced, its one-FMA filter passes the term, and the
    exact re-test must reject it.  T(n) is placed at every position of a
    30-index job (the blocks' filtered, exact and state positions, and the
    tail)."""
    T = [0, 0, 1]
    while len(T) <= n:
        T.append(T[-1] + T[-2] + T[-3])
    assert (T[n] - r) % q == 0 and T[n] % q
    f0, f1 = q - 600, q + 600
    N = 30
    lo = n - (N - 1)
    want = o.trial_factor(o.TRIBONACCI, 0, lo, n + N, f0, f1)
    for off in range(N):
        n0 = n - off
        g, _, _ = gpu_lists(engine, o.TRIBONACCI, 0, n0, n0 + N, f0, f1)
        assert q not in g[off], (q, off)
        assert g == want[n0 - lo : n0 - lo + N], (q, off)


@pytest.mark.parametrize("seq", [o.FIBONACCI, o.LUCAS])
@pytest.mark.parametrize("n0,n1,f0,f1,chunk", [
    (1, 5000, 0, 1 << 18, 0),               # small primes (hits everywhere), 5 in the window
    (3, 4100, 1 << 30, (1 << 30) + (1 << 16), 0),
    (40000, 43000, (1 << 42) + (1 << 30), (1 << 42) + (1 << 30) + (1 << 14), 0),
    (7, 20000, 3, 1 << 16, 3000),           # chunks start mid-range: theta^n0 jump-ahead
])
def test_theta_form_windows(engine, seq, n0, n1, f0, f1, chunk):
    kw = {"chunk_length": chunk} if chunk else {}
    g, gc, _ = gpu_lists(engine, seq, 0, n0, n1, f0, f1, **kw)
    assert_same(g, gc, seq, 0, n0, n1, f0, f1)


# default (the short step where it is exact), the shifted step, the
# exact-product step, the general step
SM_FORM_FLAGS = (0, nat.CTD_RUN_SM_NO_SHORT, nat.CTD_RUN_SM_NO_SHIFT | nat.CTD_RUN_SM_NO_SHORT,
                 nat.CTD_RUN_SM_FULL_STEP)


@pytest.mark.parametrize("base", [10, 2, 16, 3])
def test_smarandache_step_forms_with_hits(engine, base):
    # primes in [2^20, 2^21) are above 8 n1, so the FP64 form runs; p m is far
    # below 2^53, so it takes the short step by default; CTD_RUN_SM_NO_SHORT
    # runs the shifted step ((p / 2 + n1) o is far below 2^53 too), adding
    # CTD_RUN_SM_NO_SHIFT the exact-product step, and CTD_RUN_SM_FULL_STEP
    # the general step on the same primes.  About 250 hits per base.
    n0, n1 = 1, 5000
    f0, f1 = 1 << 20, 1 << 21
    r = o.trial_factor(o.SMARANDACHE, base, n0, n1, f0, f1, threads=16)
    assert sum(len(x) for x in r) > 100
    for flags in SM_FORM_FLAGS:
        for pair in (0, nat.CTD_RUN_SM_NO_PAIR):
            g, _, _ = gpu_lists(engine, o.SMARANDACHE, base, n0, n1, f0, f1, flags=flags | pair)
            assert g == r, flags | pair


@pytest.mark.parametrize("base", [10, 3, 7, 2, 16, 12])
def test_smarandache_shifted_step_segment_edges(engine, base):
    """The shifted step tests index n after the step n -> n+1, one index
    late: the first step of a segment repeats a tested index (skipped) and the
    last index of the segment is tested after the state is restored.  Each of
    the first hits of the window is placed at the start, at the second index,
    at the last index and at every position of a 16-index block of a job, and
    the jobs must equal the oracle in every step form."""
    f0, f1 = 1 << 14, 1 << 16  # above 8 n1 for n1 <= 2048: the FP64 forms run
    r = o.trial_factor(o.SMARANDACHE, base, 1, 2048, f0, f1)
    hits = [1 + i for i, x in enumerate(r) if x][:12]
    assert len(hits) >= 8
    ranges = set()
    for h in hits:
        ranges |= {(h, h + 40), (max(1, h - 1), h + 40), (max(1, h - 2), h + 1), (max(1, h - 17), h + 1)}
        ranges |= {(max(1, h - 1 - j), h + 19) for j in (0, 1, 7, 14, 15, 16)}
    for n0, n1 in sorted(ranges):
        want = o.trial_factor(o.SMARANDACHE, base, n0, n1, f0, f1)
        for flags in SM_FORM_FLAGS + (nat.CTD_RUN_SM_NO_PAIR, nat.CTD_RUN_SM_NO_PAIR | nat.CTD_RUN_SM_NO_SHORT):
            g, _, _ = gpu_lists(engine, o.SMARANDACHE, base, n0, n1, f0, f1, flags=flags)
            assert g == want, (n0, n1, flags)


def test_smarandache_prime_dividing_the_base(engine):
    # p = 61 divides every digit multiplier of base 61: the shifted step's test
    # (S(n) m = 0 <=> S(n) = 0) needs p > b, so waves holding such a prime
    # take the other steps.  p > 8 n1 keeps the FP64 form.
    for n0, n1 in ((1, 5), (2, 7), (3, 4)):
        want = o.trial_factor(o.SMARANDACHE, 61, n0, n1, 41, 4000)
        for flags in SM_FORM_FLAGS + (nat.CTD_RUN_SM_NO_PAIR | nat.CTD_RUN_SM_NO_SHORT,):
            g, _, _ = gpu_lists(engine, o.SMARANDACHE, 61, n0, n1, 41, 4000, flags=flags)
            assert g == want, (n0, n1, flags)


def _canon(v, p):
    v %= p
    return v if v <= (p - 1) // 2 else v - p


@pytest.mark.parametrize("seq,flags", [(o.PROTH, None), (o.PROTH_FIXED_N, nat.CTD_RUN_STEPPED),
                                       (o.PROTH_FIXED_N, None)])
@pytest.mark.parametrize("j", [0, 5, 63])
def test_word_filter_collision_is_not_a_hit(engine, seq, flags, j):
    """The 64-index stride forms compare 32-bit words (the low bits of the
    residue); a state that differs from a target by exactly 2^32 matches its
    word without being a hit.  The job is built so that the first block of a
    prime p0 > 2^37 starts on such a state; the replay must reject it, so the
    factors equal the oracle's."""
    # x_n0 = T_j -+ 2^32 for the stride target T_j: equal low 32 bits
    f0, f1 = (1 << 40) + 12345, (1 << 40) + 12345 + 4096
    primes = [int(x) for x in o.sieve(f0, f1)]
    p0 = primes[len(primes) // 2]
    N = 600
    if seq == o.PROTH:
        # state x_n = k 2^n, targets T_j = -2^-(j+1)
        n0 = 1000
        t = _canon(-pow(2, -(j + 1), p0), p0)
        x = t + (1 << 32) if t < 0 else t - (1 << 32)  # same low word, still canonical
        param = (x * pow(2, -n0, p0)) % p0
        hit_index = n0 + j + 1
        assert (param * pow(2, hit_index, p0) + 1) % p0 != 0
    else:
        # state x_k = k c + 1 (c = 2^n), targets T_j = -(j+1) c
        param = 777
        c = pow(2, param, p0)
        t = _canon(-(j + 1) * c, p0)
        x = t + (1 << 32) if t < 0 else t - (1 << 32)  # same low word, still canonical
        n0 = ((x - 1) * pow(c, -1, p0)) % p0
        hit_index = n0 + j + 1
        assert (hit_index * c + 1) % p0 != 0
    g, gc, _ = gpu_lists(engine, seq, param, n0, n0 + N, f0, f1, flags=flags)
    assert p0 not in g[hit_index - n0]
    assert_same(g, gc, seq, param, n0, n0 + N, f0, f1)
