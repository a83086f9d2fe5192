"""Serving form (ctd_hip_server_*): every job's factors buffer equals the
oracle's and the plain form's, bit-exact, whether the job runs alone or
interleaved with others that arrive while it is on the device."""
from __future__ import annotations

import time

import numpy as np
import pytest

import cutrialdive_amd as ctd
from cutrialdive_amd import _native as nat
from cutrialdive_amd import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def job_of(s):
    return nat.CtdJob(s.seq_id, s.max_factors, s.param, s.n0, s.n1, s.f0, s.f1)


def assert_same(c, f, rc, rf, mx, tag):
    assert np.array_equal(np.asarray(c), np.asarray(rc)), (tag, np.nonzero(np.asarray(c) != np.asarray(rc))[0][:5])
    for i, k in enumerate(np.asarray(c).tolist()):
        if 0 < k <= mx:
            assert sorted(f[i * mx:i * mx + k].tolist()) == sorted(rf[i * mx:i * mx + k].tolist()), (tag, i)


def small_jobs(count, seed, window_bits=21, nmax=1200):
    rng = np.random.default_rng(seed)
    fams = [(0, 0), (1, 3), (2, 5), (3, 10), (3, 16), (4, 0), (5, 0), (6, 100), (7, 33), (8, 0), (3, 7)]
    bases = [0, 1 << 24, 1 << 32, (1 << 38) + (3 << 30), (1 << 42) + (1 << 30)]
    out = []
    for j in range(count):
        seq, param = fams[int(rng.integers(len(fams)))]
        n0 = int(rng.integers(2, 300)) if seq not in (6, 7) else int(rng.integers(1, 9999)) | 1
        n1 = n0 + int(rng.integers(1, nmax))
        f0 = bases[int(rng.integers(len(bases)))] + int(rng.integers(0, 1000))
        out.append(workloads.JobSpec("x", seq, param, n0, n1, f0, f0 + (1 << window_bits), 64))
    return out


@pytest.fixture(scope="module")
def server(engine):
    # the shapes of the batch autotuner are shared with the server
    srv = ctd.Server(engine, tile_length=1 << 20, streams=4)
    yield srv
    srv.close()


def test_alone_equals_in_the_mix(server):
    probe = workloads.JobSpec("x", 3, 10, 10, 900, (1 << 38) + 12345, (1 << 38) + 12345 + (1 << 22), 64)
    print("alone", flush=True)
    t = server.submit(job_of(probe))
    alone_c, alone_f, lat_alone = server.wait(t)
    others = small_jobs(40, seed=1)
    tickets = []
    for k, s in enumerate(others):
        tickets.append((s, server.submit(job_of(s))))
        if k == 20:
            tp = server.submit(job_of(probe))  # arrives while the others are on the device
        time.sleep(0.0005)
    mix_c, mix_f, lat_mix = server.wait(tp)
    assert_same(mix_c, mix_f, alone_c, alone_f, 64, "probe")
    rc, rf = oracle.trial_factor_raw(probe.seq_id, probe.param, probe.n0, probe.n1, probe.f0, probe.f1, 64)
    assert_same(mix_c, mix_f, rc, rf, 64, "probe vs oracle")
    for s, t in tickets:
        c, f, _ = server.wait(t)
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, s)
    assert lat_alone > 0 and lat_mix > 0


def test_stream_of_jobs_matches_oracle(server):
    jobs = small_jobs(120, seed=2, window_bits=20, nmax=800)
    tickets = [(s, server.submit(job_of(s))) for s in jobs]
    res = [(s,) + server.wait(t)[:2] for s, t in tickets]
    st = server.stats()
    assert st.completed >= len(jobs) and st.failed == 0
    for s, c, f in res[:40]:
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, s)


def test_resume_buffers_accumulate(server):
    s = workloads.JobSpec("x", 1, 3, 1, 500, 3, 1 << 22, 64)
    lo, hi = s.with_window(3, 1 << 21), s.with_window(1 << 21, 1 << 22)
    c = np.zeros(s.n1 - s.n0, np.uint32)
    f = np.zeros((s.n1 - s.n0) * 64, np.uint64)
    server.wait(server.submit(job_of(lo), c, f))
    server.wait(server.submit(job_of(hi), c, f))
    rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
    assert_same(c, f, rc, rf, 64, s)


def test_edge_jobs(server):
    jobs = [workloads.JobSpec("x", 0, 0, 2, 50, 2, 3, 8), workloads.JobSpec("x", 1, 3, 1, 40, 0, 2, 8),
            workloads.JobSpec("x", 3, 10, 10, 10, 2, 1000, 8), workloads.JobSpec("x", 8, 0, 2, 70, 3, 4, 8)]
    for s in jobs:
        c, f, _ = server.wait(server.submit(job_of(s)))
        if s.n1 > s.n0:
            rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 8)
            assert_same(c, f, rc, rf, 8, s)


def test_invalid_job_rejected_at_submit(server):
    with pytest.raises(RuntimeError, match="not supported"):
        server.submit(nat.CtdJob(3, 8, 99, 1, 10, 3, 100))


def test_memory_budget_and_queue_limit(engine):
    # a one-byte budget admits one job at a time; the queue holds at most two more
    big = workloads.JobSpec("x", 1, 3, 1, 100000, 1 << 40, (1 << 40) + (1 << 29), 8)
    small = workloads.JobSpec("x", 4, 0, 1, 100, 3, 1 << 16, 8)
    with ctd.Server(engine, memory_budget=1, queue_limit=2) as srv:
        tb = srv.submit(job_of(big))
        time.sleep(0.05)  # admitted: the queue is empty again
        t1 = srv.submit(job_of(small))
        t2 = srv.submit(job_of(small))
        with pytest.raises(BlockingIOError):
            srv.submit(job_of(small))
        for t in (tb, t1, t2):
            srv.wait(t)
        st = srv.stats()
        assert st.active_max == 1 and st.pending_max == 2


def test_fresh_buffers_ignore_host_contents(engine):
    # pooled device buffers are reused: a fresh job must not see an earlier job's factors
    first = workloads.JobSpec("x", 0, 0, 2, 400, 3, 1 << 22, 16)
    second = workloads.JobSpec("x", 1, 3, 2, 400, 3, 1 << 22, 16)
    with ctd.Server(engine, fresh_buffers=True) as srv:
        srv.wait(srv.submit(job_of(first)))
        junk_c = np.full(398, 7, np.uint32)
        junk_f = np.full(398 * 16, 5, np.uint64)
        c, f, _ = srv.wait(srv.submit(job_of(second), junk_c, junk_f))
    rc, rf = oracle.trial_factor_raw(second.seq_id, second.param, second.n0, second.n1, second.f0, second.f1, 16)
    assert np.array_equal(c, rc) and np.array_equal(np.sort(f.reshape(398, 16), axis=1),
                                                    np.sort(rf.reshape(398, 16), axis=1))


def test_poll_delivers_and_close_collects(engine):
    """Results reach the caller's buffers when poll reports a job done, and
    close() delivers the results of jobs nobody collected."""
    jobs = small_jobs(8, seed=3, window_bits=18, nmax=300)
    bufs = [(np.zeros(s.n1 - s.n0, np.uint32), np.zeros((s.n1 - s.n0) * 64, np.uint64)) for s in jobs]
    with ctd.Server(engine, fresh_buffers=True) as srv:
        tickets = [srv.submit(job_of(s), c, f) for s, (c, f) in zip(jobs, bufs)]
        deadline = time.time() + 60
        while srv.poll(tickets[0])[0] != 2:
            assert time.time() < deadline
            time.sleep(0.001)
        c, f = bufs[0]
        s = jobs[0]
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, "poll")
    for s, (c, f) in zip(jobs[1:], bufs[1:]):
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, "close")


def test_fifo_policy_gives_the_same_buffers(engine):
    # the tile order (by default alternately the oldest job and the least
    # remaining work, CTD_SERVE_FIFO oldest first) changes only the
    # interleaving, never a job's factors
    jobs = small_jobs(30, seed=7, window_bits=22, nmax=3000)
    out = {}
    for fifo in (False, True):
        print("fifo", fifo, flush=True)
        srv = ctd.Server(engine, tile_length=1 << 20, streams=4, fifo=fifo)
        tickets = [srv.submit(job_of(s)) for s in jobs]
        out[fifo] = [srv.wait(t)[:2] for t in tickets]
        srv.close()
    for s, (c0, f0), (c1, f1) in zip(jobs, out[False], out[True]):
        assert_same(c0, f0, c1, f1, 64, s)
    for s, (c, f) in list(zip(jobs, out[False]))[:8]:
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, s)
