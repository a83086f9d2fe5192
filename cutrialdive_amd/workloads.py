"""Seeded synthetic workloads of the batched and serving forms.

The mixed-job workload: 256 jobs, each family uniform over {proth k, riesel k,
proth_fixed_n, fibonacci, lucas, tribonacci, smarandache b in 2..16}; k odd and
log-uniform in [3, 10^7); the fixed exponent n of proth_fixed_n log-uniform in
[10^2, 10^5]; index-range lengths log-uniform in [10^2, 10^5]; index starts
log-uniform in [10^2, 10^5] (n-indexed families) or odd in [3, 10^7) (k of the
fixed-n family); prime windows 2^30-wide slices 2^38 + j 2^30, j uniform in
[0, 7936), of [2^38, 2^43).  Every index start is >= 100, so no job touches an
index where S(n) = 0 (where every prime would divide it).

The serving workload draws its jobs from the same distribution and adds
Poisson arrival times.

Python's own Mersenne Twister (random.Random) with a fixed seed makes the
lists identical on every machine.
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass
from typing import List

SEED = 20261203
KINDS = ["proth", "riesel", "proth_fixed_n", "fibonacci", "lucas", "tribonacci", "smarandache"]
SEQ_IDS = {"mersenne": 0, "proth": 1, "riesel": 2, "smarandache": 3, "fibonacci": 4, "lucas": 5,
           "proth_fixed_n": 6, "riesel_fixed_n": 7, "tribonacci": 8}
WINDOW_BASE = 1 << 38
WINDOW_BITS = 30
WINDOW_SLICES = ((1 << 43) - (1 << 38)) >> WINDOW_BITS  # 7936


@dataclass
class JobSpec:
    kind: str
    seq_id: int
    param: int
    n0: int
    n1: int
    f0: int
    f1: int
    max_factors: int = 32

    @property
    def spec(self) -> str:
        """--num-seq text of the job."""
        return self.kind if self.kind in ("fibonacci", "lucas", "tribonacci") else f"{self.kind},{self.param}"

    def with_window(self, f0: int, f1: int) -> "JobSpec":
        return JobSpec(self.kind, self.seq_id, self.param, self.n0, self.n1, f0, f1, self.max_factors)

    def with_indices(self, n0: int, n1: int) -> "JobSpec":
        return JobSpec(self.kind, self.seq_id, self.param, n0, n1, self.f0, self.f1, self.max_factors)


def _log_uniform(rng: random.Random, lo: float, hi: float) -> int:
    return int(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def _odd_log_uniform(rng: random.Random, lo: int, hi: int) -> int:
    """Odd integer in [lo, hi) (lo odd), log-uniform."""
    v = max(_log_uniform(rng, lo, hi) | 1, lo)
    return v if v < hi else (hi - 1) | 1 if (hi - 1) | 1 < hi else hi - 2


def draw_job(rng: random.Random, window_bits: int = WINDOW_BITS, max_factors: int = 32) -> JobSpec:
    kind = KINDS[rng.randrange(len(KINDS))]
    if kind in ("proth", "riesel"):
        param = _odd_log_uniform(rng, 3, 10 ** 7)
    elif kind == "proth_fixed_n":
        param = _log_uniform(rng, 1e2, 1e5)
    elif kind == "smarandache":
        param = rng.randrange(2, 17)
    else:
        param = 0
    length = _log_uniform(rng, 1e2, 1e5)
    if kind == "proth_fixed_n":
        n0 = _odd_log_uniform(rng, 3, 10 ** 7)
    else:
        n0 = _log_uniform(rng, 1e2, 1e5)
    j = rng.randrange(WINDOW_SLICES)
    f0 = WINDOW_BASE + (j << WINDOW_BITS)
    return JobSpec(kind, SEQ_IDS[kind], param, n0, n0 + length, f0, f0 + (1 << window_bits), max_factors)


def mixed_jobs(count: int = 256, seed: int = SEED, window_bits: int = WINDOW_BITS) -> List[JobSpec]:
    """The mixed-job batch (256 jobs by default)."""
    rng = random.Random(seed)
    return [draw_job(rng, window_bits) for _ in range(count)]


def poisson_arrivals(count: int, rate_per_s: float, seed: int = SEED + 1) -> List[float]:
    """Arrival times (seconds from 0) of a Poisson stream of the given rate."""
    rng = random.Random(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += rng.expovariate(rate_per_s)
        out.append(t)
    return out


def updates_of(jobs: List[JobSpec]) -> float:
    """Expected residue updates: primes of each window (x / ln x estimate) x indices."""
    tot = 0.0
    for j in jobs:
        tot += (j.f1 - j.f0) / math.log(j.f1) * (j.n1 - j.n0)
    return tot
