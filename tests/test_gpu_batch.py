"""Batched form (ctd_hip_run_batch): every job's factors buffer equals the
oracle's, bit-exact, whatever the tile length, stream count, graph or eager
issue and autotuned launch shape; the same graph-built batch run several
times in a row gives the same buffers each time."""
from __future__ import annotations

import os

import numpy as np
import pytest

from cutrialdive_amd import _native as nat
from cutrialdive_amd import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def job_of(s: workloads.JobSpec) -> nat.CtdJob:
    return nat.CtdJob(s.seq_id, s.max_factors, s.param, s.n0, s.n1, s.f0, s.f1)


def assert_same(counts, factors, ref_counts, ref_factors, mx, ref_mx, tag):
    counts = np.asarray(counts)
    assert np.array_equal(counts, ref_counts), (tag, np.nonzero(counts != ref_counts)[0][:5])
    for i, c in enumerate(counts.tolist()):
        if c == 0 or c > mx:
            continue
        got = sorted(factors[i * mx:i * mx + c].tolist())
        exp = sorted(ref_factors[i * ref_mx:i * ref_mx + c].tolist())
        assert got == exp, (tag, i, got, exp)


def oracle_of(s: workloads.JobSpec, mx=None):
    mx = mx or s.max_factors
    return oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, mx)


def small_mixed(count=20, seed=7, window=1 << 22, nmax=1500):
    """Jobs of every family on narrow windows at several prime sizes."""
    rng = np.random.default_rng(seed)
    bases = [0, 1 << 20, 1 << 32, (1 << 38) + (5 << 30), (1 << 42) + (1 << 30), (1 << 54) - (1 << 24)]
    fams = [(0, 0), (1, 3), (2, 5), (3, 10), (3, 2), (3, 16), (4, 0), (5, 0), (6, 100), (7, 33), (8, 0),
            (1, 1279), (2, 9999991), (3, 7), (6, 3), (3, 13)]
    out = []
    for j in range(count):
        seq, param = fams[j % len(fams)]
        n0 = int(rng.integers(1, 300)) if seq not in (6, 7) else int(rng.integers(1, 9999)) | 1
        if seq == 8:
            n0 = max(n0, 2)
        n1 = n0 + int(rng.integers(1, nmax))
        f0 = bases[j % len(bases)] + int(rng.integers(0, 1000))
        w = window if f0 < (1 << 50) else window >> 4
        out.append(workloads.JobSpec("x", seq, param, n0, n1, f0, f0 + w, 64))
    return out


@pytest.fixture(scope="module")
def small_ref():
    jobs = small_mixed()
    return jobs, [oracle_of(s) for s in jobs]


@pytest.mark.parametrize("flags,tile,streams", [
    (0, 0, 0),
    (0, 1 << 20, 3),
    (nat.CTD_BATCH_NO_GRAPH, 1 << 21, 4),
    (nat.CTD_BATCH_NO_AUTOTUNE, 1 << 20, 1),
    (nat.CTD_BATCH_NO_GRAPH | nat.CTD_BATCH_NO_AUTOTUNE, 0, 16),
])
def test_small_mixed_batch_matches_oracle(engine, small_ref, flags, tile, streams):
    jobs, ref = small_ref
    print("batch flags", flags, "tile", tile, "streams", streams, flush=True)
    res, st = engine.run_batch([job_of(s) for s in jobs], flags=flags, tile_length=tile, streams=streams)
    for k, (s, (c, f), (rc, rf)) in enumerate(zip(jobs, res, ref)):
        assert_same(c, f, rc, rf, s.max_factors, s.max_factors, (k, s))
    assert st.primes > 0 and st.tiles >= len(jobs)


def test_mixed_jobs_first_window_slice(engine):
    # the mixed-job batch, each job on the first 2^22 integers of its window and
    # at most 1500 indices: bit-exact with the oracle
    jobs = [s.with_window(s.f0, s.f0 + (1 << 22)).with_indices(s.n0, min(s.n1, s.n0 + 1500))
            for s in workloads.mixed_jobs(48)]
    print("mixed jobs", len(jobs), flush=True)
    res, st = engine.run_batch([job_of(s) for s in jobs])
    assert st.graphs >= 1
    for k, (s, (c, f)) in enumerate(zip(jobs, res)):
        rc, rf = oracle_of(s)
        assert_same(c, f, rc, rf, s.max_factors, s.max_factors, (k, s))


@pytest.mark.parametrize("where", ["own", "low"])
def test_mixed_jobs_full_index_ranges(engine, where):
    """All 256 jobs of the mixed batch with their full index ranges (10^2 to
    10^5 indices, starts up to 10^5), each on 2^15 integers: the first ones of
    its own prime window in [2^38, 2^43) ("own"), or [2^24, 2^24 + 2^15)
    ("low", where the FP64 forms run too and 128 of the jobs have hits, 639 in
    all).  The factors buffers equal the oracle's."""
    lo = 1 << 24
    jobs = [s.with_window(s.f0, s.f0 + (1 << 15)) if where == "own" else s.with_window(lo, lo + (1 << 15))
            for s in workloads.mixed_jobs()]
    print("mixed jobs, full index ranges,", where, "windows", flush=True)
    res, st = engine.run_batch([job_of(s) for s in jobs])
    assert st.tiles >= len(jobs)
    threads = min(16, os.cpu_count() or 1)
    hits = 0
    for k, (s, (c, f)) in enumerate(zip(jobs, res)):
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, s.max_factors, threads)
        assert_same(c, f, rc, rf, s.max_factors, s.max_factors, (k, s))
        hits += int(rc.sum())
    if where == "low":
        assert hits > 256


def test_graph_replays_agree(engine):
    jobs = small_mixed(count=12, seed=11, window=1 << 21)
    first, _ = engine.run_batch([job_of(s) for s in jobs], tile_length=1 << 20, streams=4)
    for rep in range(3):
        # eager work and synchronisation between the launches
        engine.sieve(1 << 32, (1 << 32) + 4096)
        again, _ = engine.run_batch([job_of(s) for s in jobs], tile_length=1 << 20, streams=4)
        for (c0, f0), (c1, f1), s in zip(first, again, jobs):
            assert_same(c1, f1, c0, f0, s.max_factors, s.max_factors, ("replay", rep, s))


def test_batch_equals_plain_form(engine):
    jobs = small_mixed(count=10, seed=5, window=1 << 23, nmax=3000)
    res, _ = engine.run_batch([job_of(s) for s in jobs], tile_length=1 << 21)
    for s, (c, f) in zip(jobs, res):
        pc, pf, _ = engine.run_job(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, s.max_factors)
        assert_same(c, f, pc, pf, s.max_factors, s.max_factors, s)


def test_accumulates_into_existing_buffers(engine):
    # two halves of each window into the same buffers == the whole window
    jobs = small_mixed(count=8, seed=3, window=1 << 22)
    lo = [s.with_window(s.f0, s.f0 + (1 << 21)) for s in jobs]
    hi = [s.with_window(s.f0 + (1 << 21), s.f1) for s in jobs]
    bufs = [(np.zeros(s.n1 - s.n0, np.uint32), np.zeros((s.n1 - s.n0) * s.max_factors, np.uint64)) for s in jobs]
    for part in (lo, hi):
        engine.run_batch_raw([job_of(s) for s in part], [c.ctypes.data for c, _ in bufs],
                             [f.ctypes.data for _, f in bufs], nat.CTD_BATCH_HOST_BUFFERS, 1 << 20, 2)
    for s, (c, f) in zip(jobs, bufs):
        rc, rf = oracle_of(s)
        assert_same(c, f, rc, rf, s.max_factors, s.max_factors, s)


def test_edge_jobs(engine):
    jobs = [
        workloads.JobSpec("x", 0, 0, 2, 50, 2, 3, 8),         # prime 2 only
        workloads.JobSpec("x", 1, 3, 1, 40, 0, 2, 8),         # no prime at all
        workloads.JobSpec("x", 3, 10, 10, 10, 2, 1000, 8),    # empty index range
        workloads.JobSpec("x", 4, 0, 1, 60, 2, 1 << 16, 8),
        workloads.JobSpec("x", 8, 0, 2, 70, 3, 4, 8),         # the prime 3 only
    ]
    res, _ = engine.run_batch([job_of(s) for s in jobs], tile_length=1 << 20)
    for s, (c, f) in zip(jobs, res):
        if s.n1 <= s.n0:
            continue
        rc, rf = oracle_of(s)
        assert_same(c, f, rc, rf, s.max_factors, s.max_factors, s)


def test_batch_shape_is_valid(engine):
    s = workloads.mixed_jobs(1)[0]
    b, ppl, sp = engine.batch_shape(job_of(s))
    assert b in (64, 256) and ppl in (1, 2, 4) and sp in (1, 2, 4)


def test_invalid_job_is_rejected(engine):
    with pytest.raises(RuntimeError, match="job 1"):
        engine.run_batch([nat.CtdJob(1, 8, 3, 1, 10, 3, 100), nat.CtdJob(3, 8, 99, 1, 10, 3, 100)])
