"""GPU parity at the top of the 64-bit range.

--tf-bits 64 means primes in [2, 2^64 - 1) (command_line.cpp:317-325 of the
reference), so the device sieve's last segments start within 2^20 of 2^64,
where a segment's end lo + 2^20 does not fit in 64 bits.  These tests sieve
and trial-factor windows there through every form (ctd_hip_sieve, the plain
form, the batched form with tiles that straddle 2^64 - 2^20, the serving
form) and compare with the oracle, which decides primality there by a
deterministic Miller-Rabin test.

Hits are planted because a random window near 2^64 holds none: a Proth
multiplier k = -2^-n mod q makes q divide k 2^n + 1.  For a prime q the hit
must be listed; for the composite c = 2^64 - 873 (least prime factor 361373,
so a sieve that skips the sieving numbers above 8192 keeps it) it must not.
"""
from __future__ import annotations

import numpy as np
import pytest

import cutrialdive_amd as ctd
from cutrialdive_amd import _native as nat
from oracle import oracle as o

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
P_TOP = (1 << 64) - 59            # the largest prime below 2^64
C_TRAP = (1 << 64) - 873          # 361373 x 51046...: no prime factor below 8192
N_HIT = 1000


def proth_k(q: int, n: int) -> int:
    return (-pow(2, -n, q)) % q


WINDOWS = [
    ((1 << 64) - (1 << 20) + 12345, M64),
    ((1 << 64) - (1 << 21) - (1 << 19), M64),
    ((1 << 64) - 1000, M64),
    ((1 << 64) - 400, M64),
    ((1 << 64) - (1 << 20) - 7, (1 << 64) - (1 << 20) + 5000),
]


@pytest.mark.parametrize("f0,f1", WINDOWS, ids=lambda v: f"{v:x}")
def test_sieve_top_windows(engine, f0, f1):
    print(f"[top64] sieve [{f0}, {f1})", flush=True)
    g = engine.sieve(f0, f1)
    r = o.sieve(f0, f1)
    assert np.array_equal(g, r), (len(g), len(r))
    assert C_TRAP not in set(g.tolist())
    if f1 == M64 and f0 <= P_TOP:
        assert int(g[-1]) == P_TOP


def lists(counts, factors, mx):
    return [sorted(int(x) for x in factors[i * mx:i * mx + min(int(c), mx)]) for i, c in enumerate(counts)]


def assert_job(counts, factors, seq, k, n0, n1, f0, f1, mx):
    rc, rf = o.trial_factor_raw(seq, k, n0, n1, f0, f1, mx)
    assert np.array_equal(np.asarray(counts), rc)
    assert lists(counts, factors, mx) == lists(rc, rf, mx)


JOBS = [
    # (seq, param, n0, n1): a planted prime, a planted composite, and plain families
    (o.PROTH, proth_k(P_TOP, N_HIT), N_HIT - 300, N_HIT + 300),
    (o.PROTH, proth_k(C_TRAP, N_HIT), N_HIT - 300, N_HIT + 300),
    (o.PROTH, 3, 1, 2001),
    (o.MERSENNE, 0, 2, 1500),
    (o.SMARANDACHE, 10, 1, 400),
    (o.FIBONACCI, 0, 1, 700),
]


@pytest.mark.parametrize("seq,param,n0,n1", JOBS, ids=lambda v: str(v)[:12])
def test_plain_form_top_window(engine, seq, param, n0, n1):
    f0, f1 = (1 << 64) - (1 << 20) + 12345, M64
    print(f"[top64] plain seq {seq} k {param} n [{n0}, {n1}) primes [{f0}, {f1})", flush=True)
    mx = 16
    counts, factors, last = engine.run_job(seq, param, n0, n1, f0, f1, mx)
    assert_job(counts, factors, seq, param, n0, n1, f0, f1, mx)
    if param == proth_k(P_TOP, N_HIT):
        assert lists(counts, factors, mx)[N_HIT - n0] == [P_TOP]
    assert C_TRAP not in set(factors.tolist())
    assert last == P_TOP  # the largest prime of the window


def test_plain_form_small_batches_at_the_top(engine):
    # device batches of one segment: every batch but the first starts above 2^64 - 2^20
    f0, f1 = (1 << 64) - 3 * (1 << 20), M64
    seq, k = o.PROTH, proth_k(P_TOP, N_HIT)
    print(f"[top64] plain, 2^20-integer batches, primes [{f0}, {f1})", flush=True)
    opts = nat.CtdRunOpts(1 << 20, 0, 0, 0, 0)
    counts, factors, _ = engine.run_job(seq, k, N_HIT - 50, N_HIT + 50, f0, f1, 8, opts=opts)
    assert_job(counts, factors, seq, k, N_HIT - 50, N_HIT + 50, f0, f1, 8)


def test_batched_tiles_straddle_the_top(engine):
    # tiles of 2^21 integers from 2^64 - 2^21 - 2^19: the first straddles
    # 2^64 - 2^20, the second starts 2^19 below 2^64
    f0, f1 = (1 << 64) - (1 << 21) - (1 << 19), M64
    specs = [(s, p, a, b) for s, p, a, b in JOBS]
    jobs = [nat.CtdJob(s, 16, p, a, b, f0, f1) for s, p, a, b in specs]
    print(f"[top64] batched {len(jobs)} jobs, primes [{f0}, {f1}), tiles 2^21", flush=True)
    for flags in (0, nat.CTD_BATCH_NO_GRAPH):
        res, st = engine.run_batch(jobs, flags=flags, tile_length=1 << 21, streams=2)
        assert st.tiles == 2 * len(jobs)
        for (s, p, a, b), (c, f) in zip(specs, res):
            assert_job(c, f, s, p, a, b, f0, f1, 16)
        planted = res[0]
        assert lists(*planted, 16)[N_HIT - specs[0][2]] == [P_TOP]


def test_served_jobs_at_the_top(engine):
    f0, f1 = (1 << 64) - (1 << 21) - (1 << 19), M64
    print(f"[top64] serving {len(JOBS)} jobs, primes [{f0}, {f1})", flush=True)
    with ctd.Server(engine, tile_length=1 << 21, streams=2) as srv:
        tickets = [(srv.submit(nat.CtdJob(s, 16, p, a, b, f0, f1)), (s, p, a, b)) for s, p, a, b in JOBS]
        for t, (s, p, a, b) in tickets:
            c, f, _ = srv.wait(t)
            assert_job(c, f, s, p, a, b, f0, f1, 16)
