"""The VALU-per-update model that bench.py's roofline uses (CPU only).

The model must pick, for each prime window, the loop form the kernels pick:
the Smarandache short or shifted step (5 instructions) where (p / 2 + n1) o
< 2^53 for the odd part o of the digit multiplier m, the 128-word stride for
the doubling families, the Tribonacci lazy form's 11-index blocks below 2^43
and 8-index blocks below 2^44."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import pytest  # noqa: E402

import bench  # noqa: E402


def _digits_multiplier(b, x):
    m = b
    while m <= x:
        m *= b
    return m


def test_sm_multipliers_match_brute_force():
    for b, n0, n1 in ((10, 10, 1000), (2, 1, 20), (16, 1, 2000), (7, 3, 3000), (10, 1, 2)):
        want = {}
        for n in range(n0, n1 - 1):  # steps n -> n+1 append n+1
            m = _digits_multiplier(b, n + 1)
            want[m] = want.get(m, 0) + 1
        got = dict(bench.sm_multipliers(b, n0, n1))
        assert got == want, (b, n0, n1)


def test_smarandache_shifted_and_general_windows():
    t = bench.ops_table()
    shift, full = t["loop"]["smarandache"]["fp64_shift"], t["loop"]["smarandache"]["fp64"]
    # cfg2: m <= 1000, p < 2^36: every step shifted
    assert bench.window_ops(t, "smarandache", 1 << 32, 1 << 36, 990, (10, 10, 1000)) == pytest.approx(shift)
    # base 10 at p near 2^43: m = 10^5 has odd part 5^5, (p / 2 + n1) 5^5 < 2^53: shifted
    assert bench.window_ops(t, "smarandache", 1 << 42, (1 << 42) + (1 << 30), 1000,
                            (10, 20000, 21000)) == pytest.approx(shift)
    # base 13, m = 13^4 = 28561 at p near 2^43: (p / 2 + n1) m > 2^53, every step general
    assert bench.window_ops(t, "smarandache", 1 << 43, (1 << 43) + (1 << 30), 1000,
                            (13, 3000, 4000)) == pytest.approx(full)
    # a window straddling the limit of its single multiplier lies strictly between
    lim = int((bench.SM_SHORT_LIMIT / 13 ** 4 - 4000) / bench.SM_EXACT_HALF)
    mid = bench.window_ops(t, "smarandache", lim - (1 << 30), lim + (1 << 30), 1000, (13, 3000, 4000))
    assert shift < mid < full


def test_stride_forms_by_prime_size():
    t = bench.ops_table()
    words = 138 / 128
    assert bench.window_ops(t, "proth", 1 << 30, 1 << 31) == words
    assert bench.window_ops(t, "proth", 1 << 40, 1 << 41) == words
    # Tribonacci lazy form: 3 L + 7 instructions per block of L indices, L
    # from 11 (p < 2^43) down to 5 (p < 2^48)
    for lo, L in ((40, 11), (43, 10), (44, 8), (45, 7), (46, 6), (47, 5)):
        assert bench.window_ops(t, "tribonacci", 1 << lo, 1 << (lo + 1)) == pytest.approx((3 * L + 7) / L)
    assert bench.window_ops(t, "tribonacci", 1 << 48, 1 << 49) == 11.25
    assert bench.window_ops(t, "fibonacci", 1 << 40, 1 << 41, 5000) == t["loop"]["fibonacci"]["fp64_theta"]


def test_contract_ops_follow_the_fixed_cost_table():
    # W64 above 2^31: doubling 6, fixed-n 7, Fibonacci/Lucas 7, Tribonacci 13,
    # Smarandache 40 (32 for power-of-2 bases); W32 below
    big, small = 1 << 40, 1 << 30
    assert [bench.contract_ops(s, 3, big) for s in (0, 1, 2)] == [6.0, 6.0, 6.0]
    assert [bench.contract_ops(s, 3, small) for s in (0, 1, 2)] == [4.0, 4.0, 4.0]
    assert bench.contract_ops(6, 100, big) == 7.0 and bench.contract_ops(7, 100, big) == 7.0
    assert bench.contract_ops(4, 0, big) == 7.0 and bench.contract_ops(5, 0, big) == 7.0
    assert bench.contract_ops(8, 0, big) == 13.0 and bench.contract_ops(8, 0, small) == 7.0
    assert bench.contract_ops(3, 10, big) == 40.0 and bench.contract_ops(3, 16, big) == 32.0
    assert bench.contract_ops(3, 2, big) == 32.0 and bench.contract_ops(3, 7, big) == 40.0


def test_cpu_share_respects_affinity_and_omp(monkeypatch):
    n_aff = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    n, src = bench.cpu_share()
    assert n == 1 and "OMP_NUM_THREADS 1" in src and f"host logical CPUs {os.cpu_count()}" in src
    monkeypatch.delenv("OMP_NUM_THREADS")
    n, _ = bench.cpu_share()
    assert 1 <= n <= n_aff
