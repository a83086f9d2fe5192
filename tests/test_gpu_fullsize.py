"""GPU parity at the configurations' full sizes, audited by the oracle.

A full-size run is too large for the oracle to repeat, so each run is
checked two ways: (1) every factor the GPU reports divides S(n) (the oracle's
residue of S(n) modulo that prime is 0), and (2) on random sub-windows of
2^12 integers spread over the whole prime window, the GPU's lists restricted
to the sub-window's primes equal the oracle's trial factoring of that
sub-window, index by index.  (2) catches a missed factor anywhere the
sub-windows land; (1) catches a false one anywhere.
"""
import os
import random

import numpy as np
import pytest

from cutrialdive_amd import _native as nat
from cutrialdive_amd import workloads
from oracle import oracle as o

pytestmark = pytest.mark.gpu

SUB = 1 << 12
THREADS = min(16, os.cpu_count() or 1)


def hits_by_index(counts, factors, mx):
    """{index offset: sorted factors} of the indices with factors."""
    c = np.asarray(counts).astype(np.int64)
    assert (c <= mx).all(), "a count passed max_factors: the stored factors would be arbitrary"
    f = np.asarray(factors).reshape(len(c), mx)
    return {int(i): sorted(int(x) for x in f[i, : c[i]]) for i in np.nonzero(c)[0]}


def verify_hits(seq, param, n0, hits):
    """Every reported factor p of index n divides S(n)."""
    for i, ps in hits.items():
        r = o.residues_last(seq, param, n0 + i, 1, ps)
        assert not np.any(r), (seq, param, n0 + i, ps, r)


def audit(seq, param, n0, n1, f0, f1, hits, mx, nsub, seed):
    """The GPU lists restricted to random sub-windows equal the oracle's."""
    rng = random.Random(seed)
    checked = 0
    for _ in range(nsub):
        a = rng.randrange(f0, max(f0 + 1, f1 - SUB)) & ~1
        a = max(a, f0)
        b = min(a + SUB, f1)
        rc, rf = o.trial_factor_raw(seq, param, n0, n1, a, b, mx, THREADS)
        want = hits_by_index(rc, rf, mx)
        got = {i: [p for p in ps if a <= p < b] for i, ps in hits.items()}
        got = {i: ps for i, ps in got.items() if ps}
        assert got == want, (seq, param, a, b)
        checked += o.prime_count(a, b)
    return checked


def test_cfg2_full_window(engine):
    """Smarandache base 10, n in [10, 1000), every prime in [2^32, 2^36)
    (2.7e9 primes, 2.6e12 updates)."""
    seq, param, n0, n1, f0, f1, mx = o.SMARANDACHE, 10, 10, 1000, 1 << 32, 1 << 36, 64
    print("[cfg2] full window", flush=True)
    counts, factors, _ = engine.run_job(seq, param, n0, n1, f0, f1, mx)
    hits = hits_by_index(counts, factors, mx)
    assert sum(len(v) for v in hits.values()) > 50
    verify_hits(seq, param, n0, hits)
    assert audit(seq, param, n0, n1, f0, f1, hits, mx, 512, 2) > 512 * 100


@pytest.mark.parametrize("seq,param,n0,n1,width,nsub", [
    (o.SMARANDACHE, 16, 1, 2000, 3 << 40, 256),
    (o.PROTH, 3, 1, 100000, 1 << 36, 96),
])
def test_cfg3_wide_windows(engine, seq, param, n0, n1, width, nsub):
    """Config 3's jobs with their full index ranges: Smarandache base 16 on
    its whole window [2^40, 2^42) (1.15e11 primes, 2.3e14 updates), Proth
    k = 3 on the first 2^36 integers of it (the whole window, 1.2e16
    updates, would take about 6 minutes)."""
    f0 = 1 << 40
    f1 = f0 + width
    mx = 64
    print(f"[cfg3] seq {seq} param {param} on [2^40, 2^40 + {width})", flush=True)
    counts, factors, _ = engine.run_job(seq, param, n0, n1, f0, f1, mx)
    hits = hits_by_index(counts, factors, mx)
    assert hits
    verify_hits(seq, param, n0, hits)
    audit(seq, param, n0, n1, f0, f1, hits, mx, nsub, 3)


def test_cfg4_full_batch(engine):
    """The 256 jobs of the mixed batch at full size (2^30-integer windows in
    [2^38, 2^43), 1.3e14 updates) in the batched form: every reported factor
    verified, four random sub-windows of every job audited."""
    specs = workloads.mixed_jobs()
    jobs = [nat.CtdJob(s.seq_id, s.max_factors, s.param, s.n0, s.n1, s.f0, s.f1) for s in specs]
    print("[cfg4] 256 jobs, full windows, batched", flush=True)
    res, st = engine.run_batch(jobs)
    assert st.tiles >= len(jobs)
    total = 0
    for k, (s, (c, f)) in enumerate(zip(specs, res)):
        hits = hits_by_index(c, f, s.max_factors)
        total += sum(len(v) for v in hits.values())
        verify_hits(s.seq_id, s.param, s.n0, hits)
        audit(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, hits, s.max_factors, 4, 1000 + k)
    assert total > 0
