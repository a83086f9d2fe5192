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


def test_own_queues_give_the_same_buffers(engine):
    # CTD_SERVE_OWN_QUEUES changes where the server's streams run, never a
    # job's factors
    jobs = small_jobs(24, seed=11, window_bits=22, nmax=3000)
    out = {}
    for own in (False, True):
        print("own queues", own, flush=True)
        srv = ctd.Server(engine, tile_length=1 << 20, streams=3, own_queues=own)
        tickets = [srv.submit(job_of(s)) for s in jobs]
        out[own] = [srv.wait(t)[:2] for t in tickets]
        srv.close()
    for s, (c0, f0), (c1, f1) in zip(jobs, out[False], out[True]):
        assert_same(c0, f0, c1, f1, 64, s)
    for s, (c, f) in list(zip(jobs, out[True]))[:8]:
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, 64)
        assert_same(c, f, rc, rf, 64, s)


def size_class(b):
    c = 4096
    while c < b:
        c <<= 1
    return c


def device_charge(N, mx):
    # the job's device factors buffer and accumulators, in pool size classes
    return size_class(4 * N) + size_class(8 * N * mx) + size_class(16)


def pinned_charge(N, mx):
    # its pinned staging buffer: accumulators, counts, factors
    return size_class(((16 + 4 * N + 7) & ~7) + 8 * N * mx)


def test_config5_poisson_stream(engine):
    """Config 5 at test size: a seeded Poisson stream of jobs drawn from the
    mixed-job distribution (7 families, 10^2..10^5 indices, windows in
    [2^38, 2^43)) with each window cut to 2^17 integers so that the oracle
    covers every job, offered at 0.9 of the server's capacity measured on the
    same jobs.  Every served job equals the oracle and the same job run alone."""
    import os
    import threading

    specs = workloads.mixed_jobs(120, seed=workloads.SEED + 5, window_bits=17)
    jobs = [job_of(s) for s in specs]
    threads = min(os.cpu_count() or 1, 16)
    print(f"[cfg5] oracle on {len(specs)} jobs, {threads} threads", flush=True)
    ref = [oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, s.max_factors, threads) for s in specs]
    with ctd.Server(engine, fresh_buffers=True) as srv:
        print("[cfg5] each job alone", flush=True)
        alone = [srv.wait(srv.submit(j))[:2] for j in jobs]
        print("[cfg5] capacity: every job at once", flush=True)
        t0 = time.perf_counter()
        tickets = [srv.submit(j) for j in jobs]
        for t in tickets:
            srv.wait(t)
        capacity = len(jobs) / (time.perf_counter() - t0)
        arrivals = workloads.poisson_arrivals(len(jobs), 0.9 * capacity, seed=workloads.SEED + 6)
        print(f"[cfg5] Poisson stream at 0.9 x {capacity:.0f} jobs/s", flush=True)
        out = [None] * len(jobs)
        ts = [None] * len(jobs)
        ready = threading.Semaphore(0)

        def consume():
            for k in range(len(jobs)):
                ready.acquire()
                out[k] = srv.wait(ts[k])

        th = threading.Thread(target=consume)
        th.start()
        t_start = time.perf_counter()
        for k, j in enumerate(jobs):
            dt = t_start + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ts[k] = srv.submit(j)
            ready.release()
        th.join()
        st = srv.stats()
    assert st.failed == 0 and st.completed == 3 * len(jobs)
    assert st.active_max > 1  # jobs overlapped on the device
    for s, (c, f, lat), (ac, af), (rc, rf) in zip(specs, out, alone, ref):
        assert lat > 0
        assert_same(c, f, rc, rf, s.max_factors, (s.kind, s.param, s.n0, s.n1, s.f0, "vs oracle"))
        assert_same(c, f, ac, af, s.max_factors, (s.kind, s.param, s.n0, s.n1, s.f0, "vs alone"))


def test_pool_bytes_stay_within_the_budget(engine):
    """Admission charges each job the size classes of its device and pinned
    buffers, cached buffers are freed to stay under the budget, and results
    nobody collects do not pile up in pinned memory: the pools' peaks stay
    within the budget plus one job (an idle server always admits a job)."""
    budget = 8 << 20
    rng = np.random.default_rng(11)
    specs = []
    for k in range(60):
        N = int(rng.choice([100, 900, 5000, 30000, 120000]))
        mx = int(rng.choice([4, 16, 32]))
        n0 = int(rng.integers(2, 500))
        f0 = int(rng.choice([3, 1 << 30, 1 << 40]))
        specs.append(workloads.JobSpec("x", 1, 3 + 2 * k, n0, n0 + N, f0, f0 + (1 << 16), mx))
    with ctd.Server(engine, memory_budget=budget, fresh_buffers=True, tile_length=1 << 20) as srv:
        tickets = [srv.submit(job_of(s)) for s in specs]
        deadline = time.time() + 120
        while srv.stats().completed < len(specs):  # nothing collected until every job is done
            assert time.time() < deadline
            time.sleep(0.01)
        res = [srv.wait(t)[:2] for t in tickets]
        st = srv.stats()
    dev1 = max(device_charge(s.n1 - s.n0, s.max_factors) for s in specs)
    pin1 = max(pinned_charge(s.n1 - s.n0, s.max_factors) for s in specs)
    print(f"[budget] device peak {st.device_bytes_peak}, pinned peak {st.pinned_bytes_peak}, "
          f"one job at most {dev1} / {pin1}, scheduler {st.scheduler_cpu_ms:.1f} ms CPU in "
          f"{st.scheduler_passes} passes", flush=True)
    assert 0 < st.device_bytes_peak <= max(budget, dev1)
    assert 0 < st.pinned_bytes_peak <= budget + pin1
    assert st.scheduler_passes > 0
    for s, (c, f) in list(zip(specs, res))[::6]:
        rc, rf = oracle.trial_factor_raw(s.seq_id, s.param, s.n0, s.n1, s.f0, s.f1, s.max_factors)
        assert_same(c, f, rc, rf, s.max_factors, s)
