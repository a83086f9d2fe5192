"""The allocation-free entry point, ctd_hip_trial_factor_device_noalloc.

Its factors buffers must equal those of ctd_hip_trial_factor_device (the
entry point it is a variant of) on the same jobs, bit-exact: counts, and the
sorted stored factors of every index.  Calls run back to back, each smaller
than the one before, so contents a call leaves in the shared workspace would
show up in the next one.  Windows beyond its limit are refused before any
work.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import cutrialdive_amd as ctd
from cutrialdive_amd import _native as nat

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def run(eng, entry, job, opts=None):
    N = job.n1 - job.n0
    c = torch.zeros(max(N, 1), dtype=torch.int32, device="cuda")
    f = torch.zeros(max(N * job.max_factors, 1), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()  # the buffers are zero before the engine's stream uses them
    last = nat.u64()
    rc = getattr(eng.lib, entry)(eng.h, ctypes.byref(job), ctypes.byref(opts) if opts is not None else None,
                                 ctypes.c_void_p(c.data_ptr()), ctypes.c_void_p(f.data_ptr()), ctypes.byref(last),
                                 None)
    eng._check(rc)
    torch.cuda.synchronize()
    return c.cpu().numpy().view(np.uint32)[:N], f.cpu().numpy().view(np.uint64)[:N * job.max_factors], last.value


def assert_same(a, b, mx, tag):
    (ca, fa, la), (cb, fb, lb) = a, b
    assert np.array_equal(ca, cb), (tag, np.nonzero(ca != cb)[0][:5])
    assert la == lb, tag
    for i, k in enumerate(ca.tolist()):
        if 0 < k <= mx:
            assert sorted(fa[i * mx:i * mx + k].tolist()) == sorted(fb[i * mx:i * mx + k].tolist()), (tag, i)


# largest first: each call is smaller than the one before it
JOBS = [
    # smarandache,10 over 4 batches of 2^25 integers above 2^32 (about 6 M primes)
    ("sm10 2^32", nat.CtdJob(3, 64, 10, 10, 1000, 1 << 32, (1 << 32) + (1 << 27))),
    # fibonacci near 2^47 with index chunks
    ("fib 2^47", nat.CtdJob(4, 16, 0, 500, 20000, (1 << 47) - (1 << 24), (1 << 47) + (1 << 23))),
    # proth 1279 from 0: prime 2 and the reference's own test values
    ("proth1279", nat.CtdJob(1, 128, 1279, 0, 300, 0, 1 << 22)),
    # a window up to the limit itself
    ("mersenne top", nat.CtdJob(0, 8, 0, 2, 200, (1 << 48) - (1 << 21), 1 << 48)),
    ("riesel tiny", nat.CtdJob(2, 8, 5, 1, 40, 3, 1000)),
]


def test_noalloc_equals_device_form_on_consecutive_calls(engine):
    for tag, job in JOBS:
        print(f"[noalloc] {tag}: seq {job.seq_id} n [{job.n0}, {job.n1}) primes [{job.f0}, {job.f1})", flush=True)
        want = run(engine, "ctd_hip_trial_factor_device", job)
        got = run(engine, "ctd_hip_trial_factor_device_noalloc", job)
        assert_same(got, want, job.max_factors, tag)
        if tag == "proth1279":
            res = ctd.FactoringResults.from_buffer(job.n0, got[0], got[1], job.max_factors)
            assert res.factors[79] == [3, 2063]  # tests/tf_tests.cpp:88-109 of the reference


def test_noalloc_twice_in_a_row_second_smaller(engine):
    big = nat.CtdJob(1, 32, 3, 1, 5000, 1 << 40, (1 << 40) + (1 << 26))
    small = nat.CtdJob(5, 32, 0, 100, 900, 1 << 36, (1 << 36) + (1 << 20))
    print("[noalloc] two calls in a row, the second smaller", flush=True)
    got_big = run(engine, "ctd_hip_trial_factor_device_noalloc", big)
    got_small = run(engine, "ctd_hip_trial_factor_device_noalloc", small)
    assert_same(got_big, run(engine, "ctd_hip_trial_factor_device", big), 32, "big")
    assert_same(got_small, run(engine, "ctd_hip_trial_factor_device", small), 32, "small")


def test_noalloc_batch_length_is_capped(engine):
    job = nat.CtdJob(6, 32, 1000, 3, 2001, 1 << 38, (1 << 38) + (1 << 27))
    opts = nat.CtdRunOpts(1 << 31, 0, 0, 0, 0)
    print("[noalloc] batch_length 2^31 asked", flush=True)
    want = run(engine, "ctd_hip_trial_factor_device", job, opts)
    got = run(engine, "ctd_hip_trial_factor_device_noalloc", job, opts)
    st = engine.last_stats()
    assert_same(got, want, 32, "fixed-n")
    assert st.sieve_launches == 4  # 2^27 integers in batches of 2^25


@pytest.mark.parametrize("f1", [(1 << 48) + 1, (1 << 63), (1 << 64) - 1])
def test_noalloc_refuses_windows_beyond_its_limit(engine, f1):
    job = nat.CtdJob(1, 8, 3, 1, 10, f1 - 1000, f1)
    with pytest.raises(RuntimeError, match="2\\^48"):
        run(engine, "ctd_hip_trial_factor_device_noalloc", job)
