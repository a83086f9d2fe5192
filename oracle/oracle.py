"""ctypes binding of the oracle (oracle/libctd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results against.  The product path never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libctd_oracle.so")

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64

MERSENNE, PROTH, RIESEL, SMARANDACHE, FIBONACCI, LUCAS, PROTH_FIXED_N, RIESEL_FIXED_N, TRIBONACCI = range(9)
PEANO = 255


class Job(ctypes.Structure):
    _fields_ = [
        ("seq", u32),
        ("max_factors", u32),
        ("param", u64),
        ("n0", u64),
        ("n1", u64),
        ("f0", u64),
        ("f1", u64),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P64 = ctypes.POINTER(u64)
        sig = {
            "ctdo_mu64": (u64, [u64]),
            "ctdo_mu128": (None, [u64, P64]),
            "ctdo_mul128hi": (None, [P64, P64, P64]),
            "ctdo_mod": (u64, [u64, u64, u64]),
            "ctdo_mulmod": (u64, [u64, u64, u64]),
            "ctdo_modpow": (u64, [u64, u64, u64]),
            "ctdo_reciprocal": (u64, [u64]),
            "ctdo_mod2by1": (u64, [u64, u64, u64, u64]),
            "ctdo_modnby1": (u64, [P64, u64, u64]),
            "ctdo_digits_in_base": (u64, [u64, u64]),
            "ctdo_sieve": (u64, [u64, u64, P64, u64]),
            "ctdo_prime_count": (u64, [u64, u64]),
            "ctdo_value_mod_2": (u64, [u32, u64, u64]),
            "ctdo_next_value_mod_2": (u64, [u32, u64, u64, u64]),
            "ctdo_residues": (ctypes.c_int, [u32, u64, u64, u64, u64, P64]),
            "ctdo_smarandache_value": (u64, [u64, u64, P64, u64]),
            "ctdo_next": (u64, [u32, u64, u64, u64, u64]),
            "ctdo_residues_last": (ctypes.c_int, [u32, u64, u64, u64, P64, u64, P64]),
            "ctdo_state": (ctypes.c_int, [u32, u64, u64, u64, P64]),
            "ctdo_trial_factor": (ctypes.c_int, [ctypes.POINTER(Job), u32, ctypes.c_void_p, ctypes.c_void_p]),
            "ctdo_job_primes": (u64, [ctypes.POINTER(Job)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p64(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(u64))


def sieve(start: int, end: int) -> np.ndarray:
    L = lib()
    n = L.ctdo_sieve(start, end, None, 0)
    out = np.zeros(n, dtype=np.uint64)
    if n:
        L.ctdo_sieve(start, end, _p64(out), n)
    return out


def prime_count(start: int, end: int) -> int:
    return int(lib().ctdo_prime_count(start, end))


def residues(seq: int, param: int, n0: int, count: int, d: int) -> List[int]:
    out = np.zeros(count, dtype=np.uint64)
    rc = lib().ctdo_residues(seq, param, n0, count, d, _p64(out))
    if rc != 0:
        raise ValueError("invalid residues request")
    return [int(x) for x in out]


def residues_last(seq: int, param: int, n0: int, count: int, moduli) -> np.ndarray:
    """S(n0+count-1) mod each modulus, by propagation from n0."""
    d = np.ascontiguousarray(np.asarray(moduli, dtype=np.uint64))
    out = np.zeros(len(d), dtype=np.uint64)
    if lib().ctdo_residues_last(seq, param, n0, count, _p64(d), len(d), _p64(out)) != 0:
        raise ValueError("invalid residues request")
    return out


def next_residue(seq: int, param: int, r: int, n: int, d: int) -> int:
    return int(lib().ctdo_next(seq, param, r, n, d))


def state(seq: int, param: int, n: int, d: int) -> List[int]:
    out = np.zeros(3, dtype=np.uint64)
    k = lib().ctdo_state(seq, param, n, d, _p64(out))
    return [int(x) for x in out[:k]]


def smarandache_value(base: int, n: int) -> int:
    L = lib()
    k = L.ctdo_smarandache_value(base, n, None, 0)
    limbs = np.zeros(max(k, 1), dtype=np.uint64)
    L.ctdo_smarandache_value(base, n, _p64(limbs), k)
    v = 0
    for i in range(k - 1, -1, -1):
        v = (v << 64) | int(limbs[i])
    return v


def trial_factor_raw(seq, param, n0, n1, f0, f1, max_factors=1024, threads=8) -> Tuple[np.ndarray, np.ndarray]:
    """counts[N], factors[N*max_factors] exactly as the factors buffer holds them."""
    N = max(n1 - n0, 0)
    counts = np.zeros(N, dtype=np.uint32)
    factors = np.zeros(max(N * max_factors, 1), dtype=np.uint64)
    job = Job(seq, max_factors, param, n0, n1, f0, f1)
    rc = lib().ctdo_trial_factor(ctypes.byref(job), threads, counts.ctypes.data, factors.ctypes.data)
    if rc != 0:
        raise ValueError("invalid job")
    return counts, factors


def trial_factor(seq, param, n0, n1, f0, f1, max_factors=1024, threads=8) -> List[List[int]]:
    """Sorted factor list per index (factors_buffer::to_factoring_results)."""
    counts, factors = trial_factor_raw(seq, param, n0, n1, f0, f1, max_factors, threads)
    f2 = factors[: len(counts) * max_factors].reshape(len(counts), max_factors) if len(counts) else factors[:0]
    return [sorted(int(x) for x in f2[i, : min(int(c), max_factors)]) for i, c in enumerate(counts)]


def job_primes(seq, param, n0, n1, f0, f1) -> int:
    job = Job(seq, 1, param, n0, n1, f0, f1)
    return int(lib().ctdo_job_primes(ctypes.byref(job)))
