"""Python host-side mirror of cutrialdive's trial-factoring interface.

Names follow the reference: num_seq_id / num_seq_spec (num_seq_id.hpp:16-27,
num_seq_spec.hpp:16-22), the parameter parsing of dispatch_num_seq
(num_seq_dispatch.hpp:119-168), trial_factoring_options
(trial_factoring_options.hpp:25-39), trial_factor (trial_factoring.hpp:20-26)
and the "<index> - <factors>" writer (factoring_results.hpp:183-201).
Every computation goes through libctd_hip.so; nothing here factors anything.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _native as nat


class NumSeqId(enum.IntEnum):
    """Persisted order of num_seq_id (num_seq_id.hpp:16-27)."""

    mersenne = 0
    proth = 1
    riesel = 2
    smarandache = 3
    fibonacci = 4
    lucas = 5
    proth_fixed_n = 6
    riesel_fixed_n = 7
    tribonacci = 8


PEANO_SEQ_ID = 255  # test-only sequence S(n) = n (tests/peano.hpp)


def num_seq_id_from_string(s: str) -> NumSeqId:
    """num_seq_id_from_string (num_seq_id.cpp:25-43)."""
    try:
        return NumSeqId[s]
    except KeyError:
        raise ValueError(f"Invalid number sequence id `{s}'") from None


@dataclass
class NumSeqSpec:
    seq_id: int
    seq_params: str = ""

    @staticmethod
    def parse(text: str) -> "NumSeqSpec":
        """--num-seq value: "<id>[,<params>]" (command_line.cpp:122-146)."""
        if "," in text:
            name, params = text.split(",", 1)
            if not params:
                raise ValueError("Invalid value for option --num-seq")
            return NumSeqSpec(num_seq_id_from_string(name), params)
        return NumSeqSpec(num_seq_id_from_string(text), "")


def _parse_u64(s: str) -> Optional[int]:
    # std::from_chars on uint64_t: leading digits, no sign, no whitespace
    digits = ""
    for ch in s:
        if ch.isdigit():
            digits += ch
        else:
            break
    if not digits:
        return None
    v = int(digits)
    return v if v < (1 << 64) else None


def spec_param(spec: NumSeqSpec) -> int:
    """The numeric parameter dispatch_num_seq parses (num_seq_dispatch.hpp:119-168)."""
    sid = spec.seq_id
    if sid in (NumSeqId.proth, NumSeqId.riesel):
        k = _parse_u64(spec.seq_params)
        if k is None:
            what = "Proth number spec" if sid == NumSeqId.proth else "Riesel sequence spec"
            raise RuntimeError(f"Missing or invalid k value `{spec.seq_params}' in {what}")
        return k
    if sid in (NumSeqId.proth_fixed_n, NumSeqId.riesel_fixed_n):
        n = _parse_u64(spec.seq_params)
        if n is None:
            what = "Proth number spec" if sid == NumSeqId.proth_fixed_n else "Riesel sequence spec"
            raise RuntimeError(f"Missing or invalid n value `{spec.seq_params}' in {what}")
        return n
    if sid == NumSeqId.smarandache:
        b = 10 if spec.seq_params == "" else _parse_u64(spec.seq_params)
        if b is None:
            raise RuntimeError(f"Invalid base `{spec.seq_params}' for Smarandache numbers")
        if not 2 <= b <= 64:
            raise RuntimeError(f"Base `{b}' is not supported for Smarandache numbers")
        return b
    return 0


@dataclass
class TrialFactoringOptions:
    n0: int
    n1: int
    f0: int
    f1: int
    output_path: Optional[str] = None
    max_factors_per_number: int = 32


@dataclass
class FactoringResults:
    """factoring_results<uint64_t, uint32_t> (factoring_results.hpp:149-201)."""

    n0: int
    factors: List[List[int]] = field(default_factory=list)
    counts: List[int] = field(default_factory=list)

    @staticmethod
    def from_buffer(n0: int, counts: np.ndarray, factors: np.ndarray, max_factors: int) -> "FactoringResults":
        """factors_buffer::to_factoring_results (factors_buffer.hpp:123-138)."""
        r = FactoringResults(n0)
        f2 = factors.reshape(len(counts), max_factors)
        for i, c in enumerate(counts.tolist()):
            k = min(c, max_factors)
            r.factors.append(sorted(int(x) for x in f2[i, :k]))
            r.counts.append(int(c))
        return r

    def __len__(self) -> int:
        return len(self.factors)

    def max_factors_count(self) -> int:
        return max(self.counts, default=0)

    def factors_count(self) -> int:
        return sum(self.counts)

    def to_text(self) -> str:
        """operator<< (factoring_results.hpp:183-201): "<n> - p1, p2\\n" per index."""
        lines = [f"{self.n0 + i} - {', '.join(str(p) for p in fs)}\n" for i, fs in enumerate(self.factors)]
        return "".join(lines)


class Engine:
    """One device, one libctd_hip engine handle."""

    def __init__(self, device: int = 0):
        self.lib = nat.load()
        h = ctypes.c_void_p()
        rc = self.lib.ctd_hip_open(device, ctypes.byref(h))
        if rc != nat.CTD_OK:
            raise RuntimeError(f"ctd_hip_open failed ({rc}): {self.lib.ctd_hip_last_error(None).decode()}")
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.ctd_hip_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != nat.CTD_OK:
            raise RuntimeError(f"libctd_hip error {rc}: {self.lib.ctd_hip_last_error(self.h).decode()}")

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        mem = nat.u64()
        cu = nat.i32()
        self._check(self.lib.ctd_hip_device_info(self.h, name, 256, ctypes.byref(mem), ctypes.byref(cu)))
        return name.value.decode(), mem.value, cu.value

    def sieve(self, start: int, end: int) -> np.ndarray:
        """Primes of [start, end) from the device sieve, ascending."""
        cnt = nat.u64()
        self._check(self.lib.ctd_hip_sieve(self.h, start, end, None, 0, ctypes.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.uint64)
        if cnt.value:
            self._check(
                self.lib.ctd_hip_sieve(
                    self.h, start, end, out.ctypes.data_as(ctypes.POINTER(nat.u64)), cnt.value, ctypes.byref(cnt)
                )
            )
        return out

    def run_job(
        self,
        seq_id: int,
        param: int,
        n0: int,
        n1: int,
        f0: int,
        f1: int,
        max_factors: int = 1024,
        counts: Optional[np.ndarray] = None,
        factors: Optional[np.ndarray] = None,
        opts: Optional[nat.CtdRunOpts] = None,
        callbacks: Optional[nat.CtdCallbacks] = None,
    ):
        """ctd_hip_trial_factor on host arrays; returns (counts, factors, last_prime)."""
        N = max(n1 - n0, 0)
        if counts is None:
            counts = np.zeros(N, dtype=np.uint32)
        if factors is None:
            factors = np.zeros(N * max_factors, dtype=np.uint64)
        job = nat.CtdJob(seq_id, max_factors, param, n0, n1, f0, f1)
        last = nat.u64()
        rc = self.lib.ctd_hip_trial_factor(
            self.h,
            ctypes.byref(job),
            ctypes.byref(opts) if opts is not None else None,
            ctypes.byref(callbacks) if callbacks is not None else None,
            counts.ctypes.data_as(ctypes.c_void_p),
            factors.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(last),
        )
        self._check(rc)
        return counts, factors, last.value

    def run_job_device(self, job: nat.CtdJob, d_counts: int, d_factors: int, stream: int = 0, opts=None) -> int:
        """ctd_hip_trial_factor_device on device pointers (HBM-resident buffers)."""
        last = nat.u64()
        rc = self.lib.ctd_hip_trial_factor_device(
            self.h,
            ctypes.byref(job),
            ctypes.byref(opts) if opts is not None else None,
            ctypes.c_void_p(d_counts),
            ctypes.c_void_p(d_factors),
            ctypes.byref(last),
            ctypes.c_void_p(stream) if stream else None,
        )
        self._check(rc)
        return last.value

    def residues(self, seq_id: int, param: int, n0: int, n1: int, moduli, flags: int = 0) -> np.ndarray:
        """Test hook: S(n1-1) mod each modulus through the residue kernels."""
        mods = np.ascontiguousarray(np.asarray(moduli, dtype=np.uint64))
        out = np.zeros(len(mods), dtype=np.uint64)
        job = nat.CtdJob(seq_id, 8, param, n0, n1, 3, 5)
        P = ctypes.POINTER(nat.u64)
        self._check(
            self.lib.ctd_hip_residues(
                self.h, ctypes.byref(job), mods.ctypes.data_as(P), len(mods), out.ctypes.data_as(P), flags
            )
        )
        return out

    def valu_probe(self, op: int, iters: int = 4096) -> float:
        """Lane-instructions per second of probe op (see _native.PROBE_OPS)."""
        rate = ctypes.c_double()
        self._check(self.lib.ctd_hip_valu_probe(self.h, op, iters, ctypes.byref(rate)))
        return rate.value

    def last_stats(self) -> nat.CtdRunStats:
        st = nat.CtdRunStats()
        self._check(self.lib.ctd_hip_last_stats(self.h, ctypes.byref(st)))
        return st

    def run_batch_raw(self, jobs, counts_ptrs, factors_ptrs, flags: int = 0, tile_length: int = 0, streams: int = 0,
                      stream: int = 0) -> nat.CtdBatchStats:
        """ctd_hip_run_batch on buffer addresses (device pointers, or host arrays with
        CTD_BATCH_HOST_BUFFERS)."""
        n = len(jobs)
        arr = (nat.CtdJob * max(n, 1))(*jobs)
        cp = (ctypes.c_void_p * max(n, 1))(*[ctypes.c_void_p(int(p)) for p in counts_ptrs])
        fp = (ctypes.c_void_p * max(n, 1))(*[ctypes.c_void_p(int(p)) for p in factors_ptrs])
        opts = nat.CtdBatchOpts(tile_length, streams, flags)
        st = nat.CtdBatchStats()
        self._check(
            self.lib.ctd_hip_run_batch(
                self.h, arr, n, cp, fp, ctypes.byref(opts), ctypes.byref(st),
                ctypes.c_void_p(stream) if stream else None,
            )
        )
        return st

    def run_batch(self, jobs, flags: int = 0, tile_length: int = 0, streams: int = 0):
        """Batched form on host arrays: returns ([(counts, factors)] per job, stats)."""
        bufs = []
        for j in jobs:
            N = max(j.n1 - j.n0, 0)
            bufs.append((np.zeros(max(N, 1), dtype=np.uint32), np.zeros(max(N * j.max_factors, 1), dtype=np.uint64)))
        st = self.run_batch_raw(
            jobs, [c.ctypes.data for c, _ in bufs], [f.ctypes.data for _, f in bufs],
            flags | nat.CTD_BATCH_HOST_BUFFERS, tile_length, streams,
        )
        out = [(c[: max(j.n1 - j.n0, 0)], f[: max(j.n1 - j.n0, 0) * j.max_factors]) for (c, f), j in zip(bufs, jobs)]
        return out, st

    def batch_shape(self, job: nat.CtdJob):
        """(block, primes per lane, chunk split) the batch autotuner picks for the job."""
        b, w, s = nat.u32(), nat.u32(), nat.u32()
        self._check(self.lib.ctd_hip_batch_shape(self.h, ctypes.byref(job), ctypes.byref(b), ctypes.byref(w),
                                                 ctypes.byref(s)))
        return b.value, w.value, s.value


def trial_factor(
    spec: NumSeqSpec,
    opts: TrialFactoringOptions,
    engine: Optional[Engine] = None,
    out: Optional[Callable[[str], None]] = None,
) -> FactoringResults:
    """trial_factor (trial_factoring.hpp:20-26) without checkpointing: returns the
    factoring results of S(n), n in [n0, n1), over the primes of [f0, f1)."""
    if opts.n1 <= opts.n0:
        if out:
            out(f"Range [{opts.n0}, {opts.n1}[ is empty, no number to process...\n")
        return FactoringResults(opts.n0)
    if opts.f1 <= opts.f0:
        if out:
            out(f"Range [{opts.f0}, {opts.f1}[ is empty, no prime numbers to divide by...\n")
        return FactoringResults(opts.n0)
    param = spec_param(spec)
    own = engine is None
    eng = engine or Engine(0)
    try:
        counts, factors, _ = eng.run_job(
            int(spec.seq_id), param, opts.n0, opts.n1, opts.f0, opts.f1, opts.max_factors_per_number
        )
    finally:
        if own:
            eng.close()
    return FactoringResults.from_buffer(opts.n0, counts, factors, opts.max_factors_per_number)


class Server:
    """Serving form (ctd_hip_server_*): submit jobs while others run.

    Each submitted job gets host numpy buffers in the factors_buffer layout;
    wait() returns them with the job's latency.  While the server is open the
    engine handle must not be used for other calls."""

    def __init__(self, engine: Engine, memory_budget: int = 0, tile_length: int = 0, streams: int = 0,
                 queue_limit: int = 0, fresh_buffers: bool = False, fifo: bool = False, depth: int = 0,
                 own_queues: bool = False):
        self.eng = engine
        self.lib = engine.lib
        flags = ((nat.CTD_SERVE_FRESH_BUFFERS if fresh_buffers else 0) | (nat.CTD_SERVE_FIFO if fifo else 0)
                 | (nat.CTD_SERVE_OWN_QUEUES if own_queues else 0))
        opts = nat.CtdServerOpts(memory_budget, tile_length, streams, queue_limit, flags, depth)
        h = ctypes.c_void_p()
        engine._check(self.lib.ctd_hip_server_open(engine.h, ctypes.byref(opts), ctypes.byref(h)))
        self.h = h
        self.jobs = {}

    def submit(self, job: nat.CtdJob, counts: Optional[np.ndarray] = None,
               factors: Optional[np.ndarray] = None) -> int:
        N = max(job.n1 - job.n0, 0)
        if counts is None:
            counts = np.zeros(max(N, 1), dtype=np.uint32)
        if factors is None:
            factors = np.zeros(max(N * job.max_factors, 1), dtype=np.uint64)
        t = nat.u64()
        rc = self.lib.ctd_hip_server_submit(self.h, ctypes.byref(job), counts.ctypes.data_as(ctypes.c_void_p),
                                            factors.ctypes.data_as(ctypes.c_void_p), ctypes.byref(t))
        if rc == nat.CTD_E_BUSY:
            raise BlockingIOError(self.lib.ctd_hip_last_error(self.eng.h).decode())
        self.eng._check(rc)
        self.jobs[t.value] = (job, counts, factors)
        return t.value

    def poll(self, ticket: int):
        st = ctypes.c_int()
        lat = ctypes.c_double()
        self.eng._check(self.lib.ctd_hip_server_poll(self.h, ticket, ctypes.byref(st), ctypes.byref(lat)))
        return st.value, lat.value

    def wait(self, ticket: int):
        """(counts, factors, latency_ms) of a finished job; its buffers are released here."""
        lat = ctypes.c_double()
        self.eng._check(self.lib.ctd_hip_server_wait(self.h, ticket, ctypes.byref(lat)))
        job, c, f = self.jobs.pop(ticket)
        N = max(job.n1 - job.n0, 0)
        return c[:N], f[:N * job.max_factors], lat.value

    def stats(self) -> nat.CtdServerStats:
        st = nat.CtdServerStats()
        self.eng._check(self.lib.ctd_hip_server_stats(self.h, ctypes.byref(st)))
        return st

    def close(self) -> None:
        if self.h:
            rc = self.lib.ctd_hip_server_close(self.h)
            self.h = None
            self.eng._check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
