#!/usr/bin/env python3
"""Prime x index residue updates/s of the MI355X trial-factoring engine.

Default workload ("mixed"): the seeded batch of 256 uneven jobs
(cutrialdive_amd/workloads.py: proth/riesel k, proth_fixed_n, fibonacci,
lucas, tribonacci, smarandache b in 2..16; 10^2..10^5 indices each; a 2^30-wide
prime window in [2^38, 2^43) each), about 1.3e14 updates.  One step runs the
whole batch in the batched form (one hipGraph launch, autotuned launch shapes,
factors buffers resident in HBM).  The same jobs then run in the plain form,
one ctd_hip_trial_factor_device call per job, for the batched-vs-plain ratio,
and the two sets of factors buffers are compared.  A bounded sample (the first
2^24 integers of each window) is also timed in the plain form with the
per-update 64-bit multiply-reduce kernels (CTD_RUN_GENERIC_ONLY), the
arithmetic shape of the reference's kernel.

Other workloads (--workload): single jobs of the plain form, one step per
device batch of 2^31 integers of the prime window:
  cfg2   Sm10, n in [10, 1000), primes in [2^32, 2^36) (30 steps = the window)
  cfg1   Mersenne, n in [2, 20000), primes < 2^30
  cfg3a  Sm16, n in [1, 2000), primes in [2^40, 2^40 + 2^36)
  cfg3b  Proth k=3, n in [1, 10^5), primes in [2^40, 2^40 + 2^32)
  readme35 / readme36 / readme37 / readme40  the reference README's benchmark
         job, Proth k = 2718281828459, n in [5e7, 5e7 + 5e5), all primes below
         2^35 / 2^36 / 2^37 / 2^40 (about 0.5 / 1 / 2 / 15 minutes);
         vs_baseline is the ratio to the reference's published rate for it

Calibration (--workload copy): one 256 MiB device-to-device copy, the
reference kernel for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes.
Probe (--workload probe): the issue rate of each VALU instruction the engine
probes (ctd_hip_valu_probe).

Serving workload (--workload serve): a Poisson stream of jobs drawn from the
same generator, with 2^24-wide prime windows so that 10^4 jobs take about a
minute, submitted to the serving form (ctd_hip_server_*) at offered loads of
0.5, 0.7 and 0.9 of its measured capacity.  Reports p50/p99 job latency and
sustained updates/s per load; every served job's factors buffer is compared
with the plain form's.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

SINGLE = {
    # name: (seq_id, param, n0, n1, F0, F1, batch, description)
    "cfg2": (3, 10, 10, 1000, 1 << 32, 1 << 36, 1 << 31,
             "smarandache,10 n in [10,1000), primes in [2^32,2^36), plain single-job form"),
    "cfg1": (0, 0, 2, 20000, 0, 1 << 30, 1 << 27,
             "mersenne n in [2,20000), primes < 2^30, plain single-job form"),
    "cfg3a": (3, 16, 1, 2000, 1 << 40, (1 << 40) + (1 << 36), 1 << 31,
              "smarandache,16 n in [1,2000), primes in [2^40, 2^40+2^36) (sub-window of [2^40,2^42))"),
    "cfg3b": (1, 3, 1, 100000, 1 << 40, (1 << 40) + (1 << 32), 1 << 28,
              "proth,3 n in [1,100000), primes in [2^40, 2^40+2^32) (sub-window of [2^40,2^42))"),
    # the reference's own benchmark job (README.md:127-141), whole TF ranges
    "readme35": (1, 2718281828459, 50000000, 50500000, 0, 1 << 35, 1 << 31,
                 "proth,2718281828459 n in [5e7, 5e7+5e5), primes < 2^35 (the reference README's benchmark job)"),
    "readme36": (1, 2718281828459, 50000000, 50500000, 0, 1 << 36, 1 << 31,
                 "proth,2718281828459 n in [5e7, 5e7+5e5), primes < 2^36 (the reference README's benchmark job)"),
    "readme37": (1, 2718281828459, 50000000, 50500000, 0, 1 << 37, 1 << 31,
                 "proth,2718281828459 n in [5e7, 5e7+5e5), primes < 2^37 (the reference README's benchmark job)"),
    "readme40": (1, 2718281828459, 50000000, 50500000, 0, 1 << 40, 1 << 31,
                 "proth,2718281828459 n in [5e7, 5e7+5e5), primes < 2^40 (the reference README's benchmark job)"),
}
# Published rates of the reference's GPU engine for the same jobs
# ((candidate x prime)/s = updates/s, GeForce RTX 5090, README.md:138-141)
PUBLISHED = {"readme35": 1.39e12, "readme36": 1.74e12, "readme37": 1.98e12, "readme40": 2.28e12}
FAMILY_NAMES = {0: "mersenne", 1: "proth", 2: "riesel", 3: "smarandache", 4: "fibonacci", 5: "lucas",
                6: "proth_fixed_n", 7: "riesel_fixed_n", 8: "tribonacci"}
METRIC = "prime×index residue updates/s, batched vs plain on same GPU, bit-exact factors"
HBM_PEAK_GBPS = 8000.0
SM_SHORT_LIMIT = 2.0 ** 53 - 2.0 ** 30  # the exactness limit of the Smarandache steps
SM_EXACT_HALF = 0.5000001  # (SM_EXACT_HALF p + n1) o below SM_SHORT_LIMIT: the shifted Smarandache step


def ops_table():
    with open(os.path.join(ROOT, "cutrialdive_amd", "ops_per_update.json")) as f:
        return json.load(f)


def loop_ops(t, fam: str, pmax: int, pmin: int, n_items: int = 0) -> float:
    """Loop VALU instructions per update of the form the residue kernel picks
    for primes in [pmin, pmax] and work items of n_items indices
    (ctd_tf_fast.hpp use_fp, ctd_tf_fp.hpp); the Smarandache FP64 form also
    needs p > 8 n1, which every bench job meets."""
    e = t["loop"][fam]
    if "closed_per_prime" in e and n_items:  # one solve per prime, whatever the index count
        return float(e["closed_per_prime"]["fp64" if pmax < (1 << 48) else "int"]) / n_items
    if pmax < (1 << 48):
        if fam in ("fibonacci", "lucas") and n_items >= 1024:
            return float(e["fp64_theta"])
        if fam in ("fibonacci", "lucas") and pmax < (1 << 31):
            return float(e["w32"])
        for bits, ops in e.get("fp64_by_bits", []):
            if pmax < (1 << bits):
                return float(ops)
        return float(e["fp64"])
    return float(e["w32"] if pmax < (1 << 31) else e["w64"])


def sm_multipliers(base: int, n0: int, n1: int):
    """[(m, steps)]: the Smarandache steps n -> n+1 over [n0, n1) grouped by
    their multiplier m = base^digits(n+1)."""
    out = []
    m, lo = base, 1  # numbers in [lo, m) have the digits of m
    while lo < n1:
        a, b = max(lo, n0 + 1), min(m, n1)
        if a < b:
            out.append((m, b - a))
        lo, m = m, m * base
    return out


def odd_part(m: int) -> int:
    return m >> ((m & -m).bit_length() - 1)


def window_ops(t, fam: str, lo: int, hi: int, n_items: int = 0, sm=None) -> float:
    """Prime-count-weighted loop ops per update over the window [lo, hi).
    sm = (base, n0, n1) for a Smarandache job: for each multiplier m its FP64
    form takes a 5-instruction step on the primes with (p / 2 + n1) o < 2^53 (o
    the odd part of m; every bench prime is above the base): the short step
    where p (0.75 m + 1) < 2^53, else the shifted step; else the general
    step."""
    cuts = {lo, hi} | {1 << b for b in (31, 35, 37, 41, 43, 44, 45, 46, 47, 48) if lo < (1 << b) < hi}
    mults = sm_multipliers(*sm) if (fam == "smarandache" and sm) else []
    lims_x = {m: int((SM_SHORT_LIMIT / odd_part(m) - sm[2]) / SM_EXACT_HALF) for m, _ in mults} if mults else {}
    cuts |= {x for x in lims_x.values() if lo < x < hi}
    cuts = sorted(cuts)
    num = den = 0.0
    for a, b in zip(cuts, cuts[1:]):
        w = (b - a) / math.log(max(b, 16))
        ops = loop_ops(t, fam, b - 1, a, n_items)
        if mults and b - 1 < (1 << 48):
            shift = float(t["loop"][fam]["fp64_shift"])
            steps = sum(c for _, c in mults)
            ops = sum(c * (shift if b <= lims_x[m] else ops) for m, c in mults) / steps
        num += w * ops
        den += w
    return num / den


def ops_per_update(workload: str, seq: int, lo: int, hi: int, n_items: int, sm=None):
    """(VALU lane-ops per update, source) for a single-job workload."""
    t = ops_table()
    m = t["measured"].get(workload)
    if m:
        return float(m["valu_lane_ops_per_update"]), "rocprofv3 " + m["profile"]
    return window_ops(t, FAMILY_NAMES[seq], max(lo, 3), hi, n_items, sm), "loop ISA count, prime-weighted"


def mixed_ops(jobs, full_batch=True):
    """Update-weighted VALU lane-ops per update of the mixed batch."""
    t = ops_table()
    m = t["measured"].get("mixed")
    if m and full_batch:
        return float(m["valu_lane_ops_per_update"]), "rocprofv3 " + m["profile"]
    num = den = 0.0
    for j in jobs:
        u = (j.f1 - j.f0) / math.log(j.f1) * (j.n1 - j.n0)
        num += u * window_ops(t, FAMILY_NAMES[j.seq_id], j.f0, j.f1, j.n1 - j.n0, (j.param, j.n0, j.n1))
        den += u
    return num / den, "loop ISA count, update-weighted"


# ctd_hip_valu_probe ops of the instructions the residue loops execute:
# v_add_f64, v_fma_f64, v_mul_f64, v_cmp_eq_f64
LOOP_PROBE_OPS = (10, 11, 12, 15)


def probe_peak(eng):
    """(probed loop peak, v_add_u32 rate) in T lane-ops/s, each the best of
    three 16384-trip probe runs (single short runs vary with the clock the
    previous work left).  The first is the fastest rate among the instructions
    the residue loops consist of, reported beside the published rate the
    roofline divides by; v_add_u32 issues faster than any of them."""
    best = lambda op: max(eng.valu_probe(op, 16384) for _ in range(3)) / 1e12
    return max(best(op) for op in LOOP_PROBE_OPS), best(0)


# Algorithmic cost contract, fixed before the first kernel was written:
# full-rate int32 lane-ops per update of each family's plain recurrence in the
# reference's form (one Barrett multiply-reduce or modular add per index), for
# primes below 2^31 (W32) and above (W64).  Independent of how this engine's
# loops compute: a form that does less work per update than the recurrence
# scores above 1 against it.
C_CONTRACT = {
    "mersenne": (4, 6), "proth": (4, 6), "riesel": (4, 6),
    "proth_fixed_n": (4, 7), "riesel_fixed_n": (4, 7),
    "fibonacci": (4, 7), "lucas": (4, 7), "tribonacci": (7, 13),
}
NOMINAL_VALU = 39.3e12  # 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz, lane-ops/s
PEAK_SOURCE = ("MI355X's published FP64 vector rate, 78.6 TFLOP/s = 39.3 T FMA lane-ops/s (256 CUs x 4 SIMD x "
               "16 lanes x 2.4 GHz); every loop instruction (v_add_f64, v_fma_f64, v_mul_f64, v_cmp_eq_f64) issues "
               "at this rate")
PROBED_SOURCE = ("fastest of the loops' instructions (v_add_f64, v_fma_f64, v_mul_f64, v_cmp_eq_f64), "
                 "independent-chain probe, best of 3 runs in this run")


def contract_ops(seq_id: int, param: int, pmax: int) -> float:
    fam = FAMILY_NAMES.get(seq_id, "peano")
    if fam == "smarandache":  # power-of-2 bases shift instead of multiplying
        return 32.0 if param & (param - 1) == 0 else 40.0
    w32, w64 = C_CONTRACT.get(fam, (4, 6))
    return float(w32 if pmax < (1 << 31) else w64)


def cpu_share():
    """(threads, description): the CPUs this process may use -- its affinity
    set, capped by a cgroup CPU quota and by OMP_NUM_THREADS when the machine
    sets them to the job's share."""
    n = len(os.sched_getaffinity(0))
    src = [f"affinity {n}"]
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, int(q) // int(per)))
            src.append(f"cgroup quota {int(q) / int(per):g}")
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
        src.append(f"OMP_NUM_THREADS {omp}")
    return n, ", ".join(src) + f"; host logical CPUs {os.cpu_count()}"


def cpu_sample(jobs, sample_len, threads):
    """Oracle (the C restatement of the reference CPU path) on the first
    sample_len integers of every job's window."""
    from oracle import oracle as o

    t0 = time.perf_counter()
    out = []
    upd = 0
    for j in jobs:
        c, f = o.trial_factor_raw(j.seq_id, j.param, j.n0, j.n1, j.f0, j.f0 + sample_len, j.max_factors, threads)
        out.append((c, f))
        upd += o.job_primes(j.seq_id, j.param, j.n0, j.n1, j.f0, j.f0 + sample_len) * (j.n1 - j.n0)
    dt = time.perf_counter() - t0
    return out, upd, dt


def same_buffers(c, f, rc, rf, mx):
    if not np.array_equal(np.asarray(c), np.asarray(rc)):
        return False
    for i, k in enumerate(np.asarray(c).tolist()):
        if 0 < k <= mx and sorted(f[i * mx:i * mx + k].tolist()) != sorted(rf[i * mx:i * mx + k].tolist()):
            return False
    return True


class DeviceBuffers:
    """The factors buffers of a list of jobs, resident in HBM (one torch
    allocation each for counts and factors)."""

    def __init__(self, torch, jobs, mx):
        self.Ns = [j.n1 - j.n0 for j in jobs]
        self.mx = mx
        self.offs = np.concatenate([[0], np.cumsum(self.Ns)]).astype(np.int64)
        self.counts = torch.zeros(int(self.offs[-1]), dtype=torch.int32, device="cuda")
        self.factors = torch.zeros(int(self.offs[-1]) * mx, dtype=torch.int64, device="cuda")
        self.cptr = [self.counts.data_ptr() + 4 * int(self.offs[k]) for k in range(len(jobs))]
        self.fptr = [self.factors.data_ptr() + 8 * mx * int(self.offs[k]) for k in range(len(jobs))]

    def zero(self):
        self.counts.zero_()
        self.factors.zero_()

    def host(self):
        return self.counts.cpu().numpy().view(np.uint32), self.factors.cpu().numpy().view(np.uint64)

    def job(self, k, c, f):
        o = self.offs
        return c[o[k]:o[k + 1]], f[o[k] * self.mx:o[k + 1] * self.mx]


def batched_vs_plain(torch, eng, nat, cjobs, mx, flags=0, tile_length=0, streams=0, steps=1, plain=True,
                     label="", bufs=None):
    """Times the batched form (one ctd_hip_run_batch per step) and then the
    plain form (one ctd_hip_trial_factor_device per job) on the same jobs and
    HBM-resident buffers, and compares the two sets of buffers."""
    bufs = bufs or DeviceBuffers(torch, cjobs, mx)
    out = {"jobs": len(cjobs)}
    upd = primes = graphs = tiles = tuned = 0
    bms = build = tune_ms = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        if k:
            bufs.zero()
        st = eng.run_batch_raw(cjobs, bufs.cptr, bufs.fptr, flags, tile_length, streams)
        upd += st.updates
        primes += st.primes
        bms += st.total_ms
        build += st.build_ms
        graphs += st.graphs
        tiles += st.tiles
        tuned += st.tuned
        tune_ms += st.autotune_ms
        print(f"# {label} batched step {k}: {st.tiles} tiles, {st.graphs} graph(s), graph {st.total_ms:.0f} ms, "
              f"build {st.build_ms:.0f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out.update(value=upd / dt, seconds=dt, updates=upd, primes=primes, graph_ms=bms, build_ms=build, graphs=graphs,
               tiles=tiles, shapes_tuned=tuned, autotune_ms=tune_ms)
    if not plain:
        return out, bufs, None
    bc, bf = bufs.host()
    bufs.zero()
    torch.cuda.synchronize()
    p_upd = p_primes = 0
    sieve_ms = tf_ms = 0.0
    per_job = []
    t0 = time.perf_counter()
    for k, c in enumerate(cjobs):
        eng.run_job_device(c, bufs.cptr[k], bufs.fptr[k])
        st = eng.last_stats()
        p_upd += st.updates
        p_primes += st.primes
        sieve_ms += st.sieve_ms
        tf_ms += st.tf_ms
        per_job.append(st.updates)
    torch.cuda.synchronize()
    pdt = time.perf_counter() - t0
    pc, pf = bufs.host()
    match = all(same_buffers(*bufs.job(k, bc, bf), *bufs.job(k, pc, pf), mx) for k in range(len(cjobs)))
    print(f"# {label} plain form {pdt:.1f} s, identical factors: {match}", file=sys.stderr, flush=True)
    plain_res = {"value": p_upd / pdt, "seconds": pdt, "updates": p_upd, "launch_form": "one job per call",
                 "factors_identical_to_batched": bool(match), "sieve_ms": sieve_ms, "tf_ms": tf_ms,
                 "primes": p_primes}
    out["batched_vs_plain"] = out["value"] / plain_res["value"]
    return out, bufs, (plain_res, per_job)


def serve_block(args, eng, ctd, nat, jobs, cjobs, batched_digests):
    """The serving form on the mixed-job distribution (config 5, bounded):
    capacity = the 256 mixed jobs submitted at once; then a seeded Poisson
    stream of jobs drawn from the same distribution (2^30-wide windows in
    [2^38, 2^43)) at args.serve_load of that capacity.  Every served job's
    buffers are compared with the batched form's on the same jobs."""
    import threading

    from cutrialdive_amd import workloads

    mx = jobs[0].max_factors

    def serve(sub, arrivals):
        srv = ctd.Server(eng, fresh_buffers=True, streams=args.serve_streams, depth=args.serve_depth,
                         tile_length=(1 << args.serve_tile_bits) if args.serve_tile_bits else 0)
        lat, dig, tickets = [0.0] * len(sub), [None] * len(sub), [None] * len(sub)
        ready = threading.Semaphore(0)

        def consume():
            for k in range(len(sub)):
                ready.acquire()
                c, f, ms = srv.wait(tickets[k])
                lat[k] = ms
                dig[k] = buffer_digest(c, f, mx)

        th = threading.Thread(target=consume)
        th.start()
        t_start = time.perf_counter()
        for k, c in enumerate(sub):
            if arrivals is not None:
                dt = t_start + arrivals[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            tickets[k] = srv.submit(c)
            ready.release()
        th.join()
        el = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        return lat, dig, el, st

    _, cap_dig, cap_s, cap_st = serve(cjobs, None)
    capacity = len(cjobs) / cap_s
    print(f"# serving capacity on the {len(cjobs)} mixed jobs: {capacity:.1f} jobs/s "
          f"({cap_st.updates / cap_s:.3e} updates/s)", file=sys.stderr, flush=True)
    specs = workloads.mixed_jobs(args.serve_block_jobs, seed=workloads.SEED + 2)
    sjobs = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f1) for j in specs]
    arr = workloads.poisson_arrivals(len(sjobs), args.serve_load * capacity, seed=workloads.SEED + 3)
    lat, dig, el, st = serve(sjobs, arr)
    lat = np.array(lat)
    print(f"# serving stream: {len(sjobs)} jobs at load {args.serve_load}: {el:.1f} s, p50 "
          f"{np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms, {st.updates / el:.3e} updates/s",
          file=sys.stderr, flush=True)
    res, _ = eng.run_batch(sjobs)
    ref = [buffer_digest(c, f, mx) for c, f in res]
    return {
        "workload": f"Poisson stream of {len(sjobs)} jobs of the mixed-job distribution (2^30-wide windows in "
                    f"[2^38, 2^43), seed {workloads.SEED + 2}) at offered load {args.serve_load} of the measured "
                    "capacity",
        "capacity_jobs_per_s": capacity,
        "capacity_updates_per_s": cap_st.updates / cap_s,
        "capacity_sample": f"the {len(cjobs)} mixed jobs submitted at once",
        "server": {"streams": args.serve_streams, "depth": args.serve_depth or 2,
                   "tile_length": 1 << (args.serve_tile_bits or 26)},
        "capacity_identical_to_batched": cap_dig == batched_digests,
        "load": args.serve_load,
        "offered_jobs_per_s": args.serve_load * capacity,
        "jobs": len(sjobs),
        "seconds": el,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "updates_per_s": st.updates / el,
        "updates": st.updates,
        "active_max": st.active_max,
        "pending_max": st.pending_max,
        "scheduler_cpu_core_frac": st.scheduler_cpu_ms / 1e3 / el,
        "factors_identical_to_batched": dig == ref,
        "latency": "submit to results back in host memory",
    }


def run_mixed(args, torch, ctd, nat):
    from cutrialdive_amd import workloads

    jobs = workloads.mixed_jobs(args.jobs, window_bits=args.window_bits)
    if args.family:
        jobs = [j for j in jobs if j.kind == args.family]
    eng = ctd.Engine(0)
    name, mem, cu = eng.device_info()
    mx = jobs[0].max_factors
    print(f"# device {name}, {cu} CUs, {mem / 2**30:.0f} GiB; {len(jobs)} mixed jobs, "
          f"{sum(j.n1 - j.n0 for j in jobs)} indices, ~{workloads.updates_of(jobs):.3e} updates", file=sys.stderr,
          flush=True)
    cjobs = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f1) for j in jobs]
    flags = (nat.CTD_BATCH_NO_AUTOTUNE if args.no_autotune else 0) | (nat.CTD_BATCH_NO_GRAPH if args.no_graph else 0)
    tile = args.tile_bits and 1 << args.tile_bits
    bufs = DeviceBuffers(torch, cjobs, mx)

    # warm-up: launch-shape probes for every (family, kind, index class), then
    # a batch on the first 2^24 integers of every window
    t0 = time.perf_counter()
    warm = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f0 + (1 << 24)) for j in jobs]
    for _ in range(args.warmup):
        eng.run_batch_raw(warm, bufs.cptr, bufs.fptr, flags, tile)
    tune_s = time.perf_counter() - t0
    shapes = {}
    if not args.no_autotune:
        for j, c in zip(jobs, cjobs):
            key = f"{j.kind}/N{'<1k' if j.n1 - j.n0 < 1024 else '<16k' if j.n1 - j.n0 < 16384 else '>=16k'}"
            shapes.setdefault(key, eng.batch_shape(c))
    bufs.zero()
    print(f"# warm-up + autotune {tune_s:.1f} s", file=sys.stderr, flush=True)

    # the timed step: the batched form; then the plain form on the same jobs
    bat, bufs, pl = batched_vs_plain(torch, eng, nat, cjobs, mx, flags, tile, args.batch_streams, args.steps,
                                     plain=not args.no_plain, label="mixed", bufs=bufs)
    value, upd, primes = bat["value"], bat["updates"], bat["primes"]
    graphs = max(bat["graphs"], 1)
    result_plain = None
    sieve = None
    c_contract = None
    if pl:
        result_plain, per_job = pl
        # the plain form's sieve launches are timed on their own (HIP events)
        sms = result_plain["sieve_ms"]
        sieve = {"primes_per_s": result_plain["primes"] / (sms / 1e3), "sieve_ms": sms,
                 "prime_stream_GBps": 8.0 * result_plain["primes"] / (sms / 1e3) / 1e9,
                 "form": "plain form, sieve_kernel launches of the 256 windows (2^31-integer batches), HIP events; "
                         "8 B per prime written to HBM and read once by the residue kernel"}
        w = np.array(per_job, dtype=np.float64)
        cs = np.array([contract_ops(c.seq_id, c.param, c.f1 - 1) for c in cjobs])
        c_contract = float((w * cs).sum() / w.sum())
        # the plain form with the reference kernel's arithmetic shape: one
        # 64-bit modular multiply-reduce per update (CTD_RUN_GENERIC_ONLY),
        # one job per call, on the first 2^generic_sample_bits integers of each
        # window (a bounded sample: this form is an order of magnitude slower)
        if args.generic_sample_bits:
            opts = nat.CtdRunOpts(0, 0, 0, 0, nat.CTD_RUN_GENERIC_ONLY)
            bufs.zero()
            torch.cuda.synchronize()
            g_upd = 0
            t0 = time.perf_counter()
            for k, c in enumerate(cjobs):
                wj = nat.CtdJob(c.seq_id, c.max_factors, c.param, c.n0, c.n1, c.f0,
                                min(c.f1, c.f0 + (1 << args.generic_sample_bits)))
                eng.run_job_device(wj, bufs.cptr[k], bufs.fptr[k], opts=opts)
                g_upd += eng.last_stats().updates
            torch.cuda.synchronize()
            gdt = time.perf_counter() - t0
            result_plain["generic_sample"] = {
                "value": g_upd / gdt, "seconds": gdt, "updates": g_upd,
                "form": "plain, one job per call, one 64-bit Montgomery multiply-reduce per update "
                        "(CTD_RUN_GENERIC_ONLY), device sieve",
                "sample": f"first 2^{args.generic_sample_bits} integers of each of the {len(cjobs)} windows"}
            result_plain["batched_vs_generic_plain"] = value / (g_upd / gdt)
            print(f"# generic plain form on the sample: {g_upd / gdt:.3e} updates/s ({gdt:.1f} s)", file=sys.stderr,
                  flush=True)
    batched_digests = None
    if args.serve_block_jobs:
        bc, bf = None, None
        # the batched form's buffers of the timed step, for the serving comparison
        bufs.zero()
        eng.run_batch_raw(cjobs, bufs.cptr, bufs.fptr, flags, tile, args.batch_streams)
        bc, bf = bufs.host()
        batched_digests = [buffer_digest(*bufs.job(k, bc, bf), mx) for k in range(len(cjobs))]
    del bufs
    torch.cuda.empty_cache()

    # short jobs: the same distribution on 2^24- and 2^20-integer windows,
    # where one job no longer fills the device
    def window_block(bits):
        sj = workloads.mixed_jobs(args.short_jobs, window_bits=bits)
        sc = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f1) for j in sj]
        # warm-up, as for the main batch: the launch shapes of every (family,
        # form, index class) of these jobs are tuned before the timed step (on
        # the first 2^20 integers of each window)
        eng.run_batch([nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f0 + (1 << 20)) for j in sj])
        sb, sbufs, (sp, _) = batched_vs_plain(torch, eng, nat, sc, mx, flags, tile, args.batch_streams, 1,
                                              label=f"2^{bits}-window mix")
        del sbufs
        torch.cuda.empty_cache()
        return {"workload": f"{len(sc)} jobs of the mixed-job distribution with 2^{bits}-wide prime windows",
                "batched": {k: sb[k] for k in ("value", "seconds", "updates", "tiles", "graphs", "shapes_tuned")},
                "plain": {k: sp[k] for k in ("value", "seconds", "updates", "factors_identical_to_batched")},
                "batched_vs_plain": sb["batched_vs_plain"]}

    short = tiny = None
    if args.short_jobs and not args.no_plain:
        short = window_block(24)
        tiny = window_block(20)

    serving = serve_block(args, eng, ctd, nat, jobs, cjobs, batched_digests) if args.serve_block_jobs else None

    c_upd, c_src = mixed_ops(jobs, full_batch=args.family is None and args.jobs == 256)
    peak, peak_add = probe_peak(eng)
    per_launch_s = bat["graph_ms"] / 1e3 / graphs
    upd_launch = upd / graphs
    achieved = upd_launch * c_upd / per_launch_s / 1e12
    sm = ops_table()["measured"]
    # the sieve's measured write and the residue kernel's measured read per prime
    traffic = (sm["sieve"]["hbm_write_bytes_per_prime"] + sm["sieve"]["hbm_read_bytes_per_prime"]
               + sm["cfg2"]["hbm_read_bytes_per_prime"]) * primes / graphs
    upd_rate_launch = upd_launch / per_launch_s
    roof = {
        "bound": "valu",
        "achieved": achieved,
        "peak": NOMINAL_VALU / 1e12,
        "unit": "T VALU lane-ops/s",
        "frac": achieved / (NOMINAL_VALU / 1e12),
        "frac_basis": "VALU issue: updates/s x the loops' VALU instructions per update (ops_per_update) / the "
                      "published issue rate of those instructions",
        "traffic": traffic,
        "traffic_unit": "HBM bytes per graph launch (prime stream: sieve write + residue read)",
        "traffic_source": "per-prime bytes measured with rocprofv3 FETCH_SIZE x 2.0 / WRITE_SIZE on the "
                          "same kernels, profiles/tf_fast_smarandache_cfg2.md",
        "kernel": "batch graph (sieve_kernel + tf_fast_kernel<family> per tile)",
        "ops_per_update": c_upd,
        "ops_per_update_source": c_src,
        "peak_source": PEAK_SOURCE,
        "probed_peak": peak,
        "frac_of_probed_peak": achieved / peak,
        "probed_peak_source": PROBED_SOURCE,
        "v_add_u32_peak": peak_add,
        "frac_of_v_add_u32_peak": achieved / peak_add,
        "frac_of_v_add_u32_peak_basis": "the same achieved rate against the probed v_add_u32 rate of this run, "
                                        "the VALU integer peak the >= 60 % target is quoted against",
        "updates_per_s_per_launch": upd_rate_launch,
        "one_compare_frac": upd_rate_launch / NOMINAL_VALU,
        "one_compare_basis": "updates/s / 39.3e12: every update costs at least one compare issued at the nominal "
                             "rate",
        "avg_launch_ms": per_launch_s * 1e3,
        "hbm": {"achieved_GBps": traffic / per_launch_s / 1e9, "peak_GBps": HBM_PEAK_GBPS,
                "frac": traffic / per_launch_s / 1e9 / HBM_PEAK_GBPS},
    }
    if c_contract is not None:
        roof["algorithmic_ops_per_update"] = c_contract
        roof["algorithmic_frac"] = upd_rate_launch * c_contract / NOMINAL_VALU
        roof["algorithmic_basis"] = ("updates/s x the fixed cost contract (int32 lane-ops per update of each "
                                     "family's plain Barrett recurrence, W64 here: doubling 6, fixed-n 7, "
                                     "Fibonacci/Lucas 7, Tribonacci 13, Smarandache 40 / 32 for power-of-2 bases), "
                                     "update-weighted over the jobs / 39.3e12; above 1 where a form replaces the "
                                     "per-update recurrence by fewer operations")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bat["seconds"] * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "dtype_note": "residues are exact integers below 2^53 carried in FP64 registers (nothing is rounded); "
                      "u64 integer forms for primes at or above 2^48",
        "data": "synthetic (seeded job generator, seed %d)" % workloads.SEED,
        "config": {
            "workload": f"mixed batch of {len(jobs)} seeded jobs (7 families, 10^2..10^5 indices, "
                        f"2^{args.window_bits}-wide prime windows in [2^38, 2^43)), batched form, one hipGraph launch",
            "jobs": len(jobs),
            "indices": int(sum(j.n1 - j.n0 for j in jobs)),
            "primes_timed": primes,
            "updates_timed": upd,
            "tiles": bat["tiles"] // max(args.steps, 1),
            "graph_launches_per_step": bat["graphs"] // max(args.steps, 1),
            "graph_build_ms_per_step": bat["build_ms"] / max(args.steps, 1),
            "autotune": not args.no_autotune,
            "shapes_tuned_in_timed_steps": bat["shapes_tuned"],
            "shapes": {k: {"block": b, "primes_per_lane": w, "chunk_split": s} for k, (b, w, s) in sorted(shapes.items())},
        },
        "plain": result_plain,
        "batched_vs_plain": (value / result_plain["value"]) if result_plain else None,
        "short_windows": short,
        "tiny_windows": tiny,
        "sieve": sieve,
        "serving": serving,
        "roofline": roof,
    }
    if not args.no_cpu:
        threads, share = cpu_share()
        sample = 1 << args.cpu_sample_bits
        ref, cupd, cdt = cpu_sample(jobs, sample, threads)
        sjobs = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f0 + sample) for j in jobs]
        res, _ = eng.run_batch(sjobs, flags=flags)
        ok = all(same_buffers(c, f, rc, rf, mx) for (c, f), (rc, rf) in zip(res, ref))
        result["cpu_baseline"] = {
            "value": cupd / cdt,
            "unit": "updates/s",
            "cores": threads,
            "cores_source": share,
            "kind": "port",
            "sample": f"oracle (C restatement of the reference CPU path) on the first 2^{args.cpu_sample_bits} "
                      f"integers of each of the {len(jobs)} job windows ({cupd:.3e} updates), {cdt:.1f} s "
                      f"with {threads} threads",
            "gpu_matches_cpu_on_sample": bool(ok),
        }
    print(json.dumps(result), flush=True)
    eng.close()
    return 0


def buffer_digest(c, f, mx):
    """Order-free digest of a factors buffer: the counts, and the stored
    factors of each index that has any, sorted (their order depends on the
    atomics)."""
    import hashlib

    c = np.ascontiguousarray(np.asarray(c).view(np.uint32))
    h = hashlib.sha1(c.tobytes())
    nz = np.nonzero(c)[0]
    if len(nz):
        rows = np.asarray(f).view(np.uint64).reshape(len(c), mx)[nz]
        k = np.minimum(c[nz], mx).astype(np.int64)
        rows = np.where(np.arange(mx)[None, :] < k[:, None], rows, np.uint64(0))
        h.update(nz.astype(np.int64).tobytes())
        h.update(np.sort(rows, axis=1).tobytes())
    return h.hexdigest()


def run_serve(args, torch, ctd, nat):
    import threading

    from cutrialdive_amd import workloads

    specs = workloads.mixed_jobs(args.serve_jobs, seed=workloads.SEED + 2, window_bits=24)
    cjobs = [nat.CtdJob(j.seq_id, j.max_factors, j.param, j.n0, j.n1, j.f0, j.f1) for j in specs]
    mx = specs[0].max_factors
    loads = [float(x) for x in args.loads.split(",")]
    eng = ctd.Engine(0)
    name, mem, cu = eng.device_info()
    print(f"# device {name}, {cu} CUs; serving {len(specs)} jobs per load, loads {loads}", file=sys.stderr,
          flush=True)
    # launch shapes for every (family, kind, index class) of the stream, shared with the server
    t0 = time.perf_counter()
    seen = {}
    for j, c in zip(specs, cjobs):
        seen.setdefault((j.seq_id, j.n1 - j.n0 < 1024, j.n1 - j.n0 < 16384), c)
    for c in seen.values():
        eng.batch_shape(c)
    print(f"# autotune {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    def serve(jobs, arrivals):
        """Submits jobs at the arrival times (None: all at once); returns
        (latencies ms, digests, seconds from first submit to last completion, stats)."""
        srv = ctd.Server(eng, tile_length=1 << 24, streams=args.serve_streams, fresh_buffers=True,
                         fifo=args.serve_fifo, depth=args.serve_depth, own_queues=args.serve_own_queues)
        lat = [0.0] * len(jobs)
        dig = [None] * len(jobs)
        tickets = [None] * len(jobs)
        ready = threading.Semaphore(0)

        def consume():
            for k in range(len(jobs)):
                ready.acquire()
                c, f, ms = srv.wait(tickets[k])
                lat[k] = ms
                dig[k] = buffer_digest(c, f, mx)

        th = threading.Thread(target=consume)
        th.start()
        t_start = time.perf_counter()
        for k, c in enumerate(jobs):
            if arrivals is not None:
                dt = t_start + arrivals[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            tickets[k] = srv.submit(c)
            ready.release()
        th.join()
        elapsed = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        return lat, dig, elapsed, st

    # capacity: the first jobs of the stream, all submitted at once
    ncap = min(len(cjobs), args.capacity_jobs)
    _, _, cap_s, cap_st = serve(cjobs[:ncap], None)
    capacity = ncap / cap_s
    print(f"# capacity {capacity:.1f} jobs/s ({cap_st.updates / cap_s:.3e} updates/s)", file=sys.stderr, flush=True)
    # the same jobs in the batched form (one graph, factors buffers resident
    # in HBM), for comparison; the serving form also returns every job's
    # buffers to host memory
    Ns = [c.n1 - c.n0 for c in cjobs[:ncap]]
    offs = np.concatenate([[0], np.cumsum(Ns)])
    d_counts = torch.zeros(int(offs[-1]), dtype=torch.int32, device="cuda")
    d_factors = torch.zeros(int(offs[-1]) * mx, dtype=torch.int64, device="cuda")
    cptr = [d_counts.data_ptr() + 4 * int(offs[k]) for k in range(ncap)]
    fptr = [d_factors.data_ptr() + 8 * mx * int(offs[k]) for k in range(ncap)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bst = eng.run_batch_raw(cjobs[:ncap], cptr, fptr, 0, 1 << 24)
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t0
    del d_counts, d_factors
    print(f"# batched form on the same jobs: {ncap / batch_s:.1f} jobs/s ({bst.updates / batch_s:.3e} updates/s)",
          file=sys.stderr, flush=True)

    # estimated VALU work of each job (primes x indices x loop ops): latency
    # is also reported separately for the 10 % of jobs with the most work
    t_ops = ops_table()
    work = np.array([(c.f1 - c.f0) / math.log(c.f1) * (c.n1 - c.n0)
                     * window_ops(t_ops, FAMILY_NAMES[c.seq_id], c.f0, c.f1, c.n1 - c.n0, (c.param, c.n0, c.n1))
                     for c in cjobs])
    large = work >= np.percentile(work, 90)

    def pct(x):
        return {"p50_ms": float(np.percentile(x, 50)), "p99_ms": float(np.percentile(x, 99))}

    # idle streams created before the served streams: HIP places a process's
    # streams on GPU_MAX_HW_QUEUES hardware queues, and how many exist before
    # the server decides which of its streams share one
    extra = [torch.cuda.Stream() for _ in range(args.serve_extra_streams)]
    # offered rates: fractions of the measured capacity, or given in jobs/s
    rates = [float(x) for x in args.serve_rates.split(",")] if args.serve_rates else [L * capacity for L in loads]
    runs = []
    for rate in rates:
        L = rate / capacity
        arr = workloads.poisson_arrivals(len(cjobs), rate, seed=workloads.SEED + int(round(L * 1000)))
        lat, dig, el, st = serve(cjobs, arr)
        lat_np = np.array(lat)
        runs.append({"load": L, "offered_jobs_per_s": rate, "jobs": len(cjobs), "seconds": el,
                     "p50_ms": float(np.percentile(lat_np, 50)), "p99_ms": float(np.percentile(lat_np, 99)),
                     "mean_ms": float(lat_np.mean()),
                     "by_work": {"largest_10pct": pct(lat_np[large]), "other_90pct": pct(lat_np[~large])},
                     "updates_per_s": st.updates / el, "updates": st.updates,
                     "tiles": st.tiles, "active_max": st.active_max, "pending_max": st.pending_max,
                     "scheduler_cpu_core_frac": st.scheduler_cpu_ms / 1e3 / el,
                     "scheduler_passes": st.scheduler_passes,
                     "device_bytes_peak": st.device_bytes_peak, "pinned_bytes_peak": st.pinned_bytes_peak,
                     "digests": dig})
        bw = runs[-1]["by_work"]
        print(f"# load {L:.2f} ({rate:.0f} jobs/s): {el:.1f} s, p50 {runs[-1]['p50_ms']:.1f} ms, "
              f"p99 {runs[-1]['p99_ms']:.1f} ms "
              f"(largest 10 %: {bw['largest_10pct']['p50_ms']:.1f} / {bw['largest_10pct']['p99_ms']:.1f}, "
              f"others: {bw['other_90pct']['p50_ms']:.1f} / {bw['other_90pct']['p99_ms']:.1f}), "
              f"{st.updates / el:.3e} updates/s, active max {st.active_max}, scheduler "
              f"{st.scheduler_cpu_ms / 1e3 / el:.2f} of a core", file=sys.stderr, flush=True)

    # plain form, one job per call: the reference the served buffers are compared with
    t0 = time.perf_counter()
    plain = []
    for c in cjobs:
        N = c.n1 - c.n0
        pc, pf, _ = eng.run_job(c.seq_id, c.param, c.n0, c.n1, c.f0, c.f1, mx)
        plain.append(buffer_digest(pc, pf, mx))
    plain_s = time.perf_counter() - t0
    served = [r["digests"] for r in runs]
    for r in runs:
        r["factors_identical_to_plain"] = r.pop("digests") == plain
    print(f"# plain form check {plain_s:.1f} s", file=sys.stderr, flush=True)

    # CPU oracle on the whole windows of the jobs with the least work: their
    # served buffers (every load) must equal the oracle's
    cpu_check = None
    if args.serve_cpu_jobs and not args.no_cpu:
        from oracle import oracle as o

        work = sorted(range(len(specs)), key=lambda k: (specs[k].n1 - specs[k].n0) * (specs[k].f1 - specs[k].f0))
        pick = work[:args.serve_cpu_jobs]
        threads, _ = cpu_share()
        t0 = time.perf_counter()
        ok = True
        for k in pick:
            j = specs[k]
            rc, rf = o.trial_factor_raw(j.seq_id, j.param, j.n0, j.n1, j.f0, j.f1, mx, threads)
            d = buffer_digest(rc, rf, mx)
            ok = ok and all(sv[k] == d for sv in served)
        cpu_check = {"jobs": len(pick), "seconds": time.perf_counter() - t0, "threads": threads,
                     "served_equal_oracle": bool(ok),
                     "selection": "the jobs of least work (indices x window), whole windows"}
        print(f"# CPU oracle check of {len(pick)} served jobs: {ok}", file=sys.stderr, flush=True)

    top = runs[-1]
    result = {
        "metric": METRIC,
        "value": top["updates_per_s"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": len(runs),
        "warmup": 0,
        "ms_per_step": top["seconds"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "dtype_note": "residues are exact integers below 2^53 carried in FP64 registers (nothing is rounded); "
                      "u64 integer forms for primes at or above 2^48",
        "data": "synthetic (seeded job generator and Poisson arrivals)",
        "config": {
            "workload": f"serving: Poisson stream of {len(cjobs)} jobs of the mixed-job distribution with 2^24-wide "
                        f"prime windows, offered "
                        + (f"rates {rates} jobs/s" if args.serve_rates else f"load {loads} of measured capacity")
                        + f"; value = sustained updates/s at {top['offered_jobs_per_s']:.0f} jobs/s",
            "extra_streams": args.serve_extra_streams,
            "capacity_jobs_per_s": capacity,
            "capacity_updates_per_s": cap_st.updates / cap_s,
            "capacity_sample_jobs": ncap,
            "batched_form_same_jobs_per_s": ncap / batch_s,
            "batched_form_same_updates_per_s": bst.updates / batch_s,
            "tile_length": 1 << 24,
            "streams": args.serve_streams,
            "policy": "fifo" if args.serve_fifo else "alternately the oldest job and the least remaining work",
            "own_queues": args.serve_own_queues,
            "depth": args.serve_depth or 2,
        },
        "serving": runs,
        "plain_check_seconds": plain_s,
        "cpu_check": cpu_check,
    }
    print(json.dumps(result), flush=True)
    eng.close()
    return 0


def run_copy(args, torch, ctd, nat):
    """HBM counter calibration: one device-to-device copy of a known size
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this run give the factors
    that scale the counters of the engine's kernels to bytes)."""
    n = 256 << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dst.copy_(src)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"workload": "copy", "bytes_read": n, "bytes_written": n, "seconds": dt,
                      "GBps": 2 * n / dt / 1e9}), flush=True)
    return 0


def run_probe(args, torch, ctd, nat):
    """Issue rate of each probed VALU instruction (lane-ops/s, best of three
    8192-trip runs): the peaks the residue loops are measured against."""
    eng = ctd.Engine(0)
    out = {}
    for op, name in enumerate(nat.PROBE_OPS):
        out[name] = max(eng.valu_probe(op, 8192) for _ in range(3))
        print(f"# {name}: {out[name] / 1e12:.2f} T lane-ops/s", file=sys.stderr, flush=True)
    print(json.dumps({"workload": "probe", "unit": "lane-ops/s", "rates": out}), flush=True)
    eng.close()
    return 0


def run_single(args, torch, ctd, nat):
    seq, param, n0, n1, F0, F1, batch, desc = SINGLE[args.workload]
    N = n1 - n0
    mx = args.max_factors
    nbatches = (F1 - F0 + batch - 1) // batch
    # jobs with a published baseline run their whole TF range by default
    steps = args.steps if args.steps is not None else (nbatches if args.workload in PUBLISHED else min(30, nbatches))
    eng = ctd.Engine(0)
    name, mem, cu = eng.device_info()
    print(f"# device {name}, {cu} CUs, {mem / 2**30:.0f} GiB; workload {args.workload}: {desc}", file=sys.stderr,
          flush=True)
    d_counts = torch.zeros(N, dtype=torch.int32, device="cuda")
    d_factors = torch.zeros(N * mx, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()

    def run_batch(k: int):
        lo = F0 + (k % nbatches) * batch
        hi = min(lo + batch, F1)
        job = nat.CtdJob(seq, mx, param, n0, n1, lo, hi)
        eng.run_job_device(job, d_counts.data_ptr(), d_factors.data_ptr())
        return eng.last_stats()

    for k in range(args.warmup):
        run_batch(k)
    d_counts.zero_()
    d_factors.zero_()
    torch.cuda.synchronize()
    updates = primes = 0
    tf_ms = sieve_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(steps):
        st = run_batch(k)
        updates += st.updates
        primes += st.primes
        tf_ms += st.tf_ms
        sieve_ms += st.sieve_ms
        launches += st.tf_launches
        print(f"# step {k}: {st.primes} primes, tf {st.tf_ms:.1f} ms, sieve {st.sieve_ms:.1f} ms", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = updates / dt
    c_upd, c_src = ops_per_update(args.workload, seq, F0, F1, N, (param, n0, n1))
    peak, peak_add = probe_peak(eng)
    avg_launch_s = tf_ms / 1e3 / max(launches, 1)
    achieved = updates / max(launches, 1) * c_upd / avg_launch_s / 1e12
    tm = ops_table()["measured"].get(args.workload)
    if tm:  # rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE of the residue kernel, per prime
        traffic = (tm["hbm_read_bytes_per_prime"] + tm["hbm_write_bytes_per_prime"]) * primes / max(launches, 1)
        traffic_src = "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE (copy-calibrated), " + tm["profile"]
    else:
        traffic = 8.0 * primes / max(launches, 1)
        traffic_src = "algorithmic: 8 B per prime read"
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (value / PUBLISHED[args.workload]) if args.workload in PUBLISHED and steps == nbatches else None,
        "vs_baseline_note": ("published rate of the reference's GPU engine on the same job and TF range: "
                             f"{PUBLISHED[args.workload]:.3g} updates/s, GeForce RTX 5090 (README.md:138-141)")
                            if args.workload in PUBLISHED else None,
        "dtype": "f64",
        "dtype_note": "residues are exact integers below 2^53 carried in FP64 registers (nothing is rounded); "
                      "u64 integer forms for primes at or above 2^48",
        "data": "synthetic",
        "config": {"workload": desc, "sequence": FAMILY_NAMES[seq], "param": param, "n0": n0, "n1": n1,
                   "prime_window": [F0, F1], "batch_integers": batch, "primes_timed": primes,
                   "updates_timed": updates, "form": "plain"},
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": NOMINAL_VALU / 1e12,
            "unit": "T VALU lane-ops/s",
            "frac": achieved / (NOMINAL_VALU / 1e12),
            "traffic": traffic,
            "traffic_unit": "HBM bytes per launch of the residue kernel",
            "traffic_source": traffic_src,
            "kernel": f"tf_fast_kernel<{FAMILY_NAMES[seq]}>",
            "ops_per_update": c_upd,
            "ops_per_update_source": c_src,
            "peak_source": PEAK_SOURCE,
            "probed_peak": peak,
            "frac_of_probed_peak": achieved / peak,
            "v_add_u32_peak": peak_add,
            "frac_of_v_add_u32_peak": achieved / peak_add,
            "avg_launch_ms": avg_launch_s * 1e3,
            "hbm": {"achieved_GBps": traffic / avg_launch_s / 1e9, "peak_GBps": HBM_PEAK_GBPS,
                    "frac": traffic / avg_launch_s / 1e9 / HBM_PEAK_GBPS},
        },
        "sieve": {
            "primes_per_s": primes / (sieve_ms / 1e3) if sieve_ms else None,
            "prime_stream_GBps": primes * 8 / (sieve_ms / 1e3) / 1e9 if sieve_ms else None,
            "sieve_ms": sieve_ms,
            "tf_ms": tf_ms,
        },
    }
    if not args.no_cpu:
        from oracle import oracle as o

        threads, share = cpu_share()
        sample = 1 << args.cpu_sample_bits
        f0s = F0 if F0 > 2 else 3
        t0 = time.perf_counter()
        ccounts, cfactors = o.trial_factor_raw(seq, param, n0, n1, f0s, f0s + sample, 1024, threads)
        cdt = time.perf_counter() - t0
        cpr = o.job_primes(seq, param, n0, n1, f0s, f0s + sample)
        gc, gf, _ = eng.run_job(seq, param, n0, n1, f0s, f0s + sample, 1024)
        ok = same_buffers(gc, gf, ccounts, cfactors, 1024)
        result["cpu_baseline"] = {
            "value": cpr * N / cdt,
            "unit": "updates/s",
            "cores": threads,
            "cores_source": share,
            "kind": "port",
            "sample": f"oracle (C restatement of the reference CPU path) on primes in [{f0s}, {f0s + sample}) "
                      f"({cpr} primes x {N} indices), {cdt:.1f} s with {threads} threads",
            "gpu_matches_cpu_on_sample": bool(ok),
        }
    print(json.dumps(result), flush=True)
    eng.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (mixed: 1, single-job workloads: 30)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed", choices=["mixed", "serve", "copy", "probe"] + sorted(SINGLE))
    ap.add_argument("--serve-jobs", type=int, default=10000, help="serve: jobs per offered load")
    ap.add_argument("--loads", default="0.5,0.7,0.9", help="serve: offered loads (fractions of capacity)")
    ap.add_argument("--serve-rates", default="", help="serve: offered rates in jobs/s (in place of --loads)")
    ap.add_argument("--serve-extra-streams", type=int, default=0,
                    help="serve: idle streams created before the served ones (hardware-queue placement)")
    ap.add_argument("--capacity-jobs", type=int, default=2000, help="serve: jobs of the capacity measurement")
    ap.add_argument("--serve-streams", type=int, default=3,
                    help="serve and the serving block of mixed: streams (resident prime slots)")
    ap.add_argument("--serve-depth", type=int, default=0,
                    help="serve and the serving block of mixed: tiles of work each stream keeps in flight "
                         "(0 = server default, 2)")
    ap.add_argument("--serve-own-queues", action="store_true",
                    help="serve: the server's streams get hardware queues of their own (CTD_SERVE_OWN_QUEUES)")
    ap.add_argument("--serve-fifo", action="store_true",
                    help="serve: tiles go to the oldest job (default: alternately the oldest and the least remaining work)")
    ap.add_argument("--serve-cpu-jobs", type=int, default=12, help="serve: served jobs checked against the oracle")
    ap.add_argument("--jobs", type=int, default=256, help="mixed: jobs drawn from the seeded generator")
    ap.add_argument("--short-jobs", type=int, default=2000,
                    help="mixed: also compare batched and plain on this many jobs with 2^24-wide windows (0: skip)")
    ap.add_argument("--serve-block-jobs", type=int, default=600,
                    help="mixed: jobs of the bounded serving stream (config 5 distribution; 0: skip)")
    ap.add_argument("--serve-load", type=float, default=0.9, help="mixed: offered load of the serving stream")
    ap.add_argument("--serve-tile-bits", type=int, default=0,
                    help="mixed: log2 of the serving tile length (0 = the server default, 2^26)")
    ap.add_argument("--tile-bits", type=int, default=0, help="mixed: log2 of the tile length (0 = engine default)")
    ap.add_argument("--batch-streams", type=int, default=0, help="mixed: streams of the batched form (0 = engine default)")
    ap.add_argument("--window-bits", type=int, default=30, help="mixed: log2 of each job's prime window")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="mixed: issue the tiles eagerly (profilers)")
    ap.add_argument("--family", default=None, help="mixed: keep only the jobs of one family (per-family rates)")
    ap.add_argument("--no-plain", action="store_true", help="mixed: skip the plain-form comparison")
    ap.add_argument("--generic-sample-bits", type=int, default=24,
                    help="mixed: also time the plain form with per-update 64-bit multiply-reduce kernels on the "
                         "first 2^B integers of each window (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample-bits", type=int, default=None)
    ap.add_argument("--max-factors", type=int, default=1024)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); the engine shards by prime "
              "window across GPUs with no exchange in the hot loop; this engine runs on one GPU.",
              file=sys.stderr)
        return 2

    import torch

    import cutrialdive_amd as ctd
    from cutrialdive_amd import _native as nat

    if args.workload == "serve":
        return run_serve(args, torch, ctd, nat)
    if args.workload == "copy":
        return run_copy(args, torch, ctd, nat)
    if args.workload == "probe":
        return run_probe(args, torch, ctd, nat)
    if args.workload == "mixed":
        if args.steps is None:
            args.steps = 1
        if args.cpu_sample_bits is None:
            args.cpu_sample_bits = 16
        return run_mixed(args, torch, ctd, nat)
    if args.cpu_sample_bits is None:
        # about 10-30 s of oracle work on 16 threads
        args.cpu_sample_bits = {"cfg1": 26, "cfg2": 29, "cfg3a": 24, "cfg3b": 20}.get(args.workload, 17)
    return run_single(args, torch, ctd, nat)


if __name__ == "__main__":
    sys.exit(main())
