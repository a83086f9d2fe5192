"""Writes trib_large_factors.json: primes 2^31 <= q < 2^48 dividing T(n) for
60 <= n < 700 (T(0) = T(1) = 0, T(2) = 1, T(n+3) = T(n+2) + T(n+1) + T(n)),
found with sympy's elliptic-curve factoring (this container only).

Each pair (n, q) is a fact about Tribonacci numbers, checked here with Python
integers (T(n) % q == 0).  The GPU tests use them to plant hits at large
primes, where a random prime window holds none.
"""
import json
import os
import time

import sympy
from sympy.ntheory import ecm, isprime


def main(budget_s=240.0):
    t0 = time.time()
    T = [0, 0, 1]
    while len(T) < 700:
        T.append(T[-1] + T[-2] + T[-3])
    found = set()
    for n in range(60, 700):
        if time.time() - t0 > budget_s:
            break
        v = T[n]
        for q in sympy.primerange(2, 20000):
            while v % q == 0:
                v //= q
        if v == 1:
            continue
        try:
            fs = {v} if isprime(v) else ecm(v, B1=20000, B2=2000000, max_curve=60, seed=n)
        except ValueError:  # no split within these bounds: skip this n
            continue
        for q in fs:
            if isprime(q) and (1 << 31) <= q < (1 << 48):
                found.add((n, int(q)))
    pairs = sorted(found, key=lambda z: z[1])
    for n, q in pairs:
        assert T[n] % q == 0
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trib_large_factors.json")
    with open(out, "w") as f:
        json.dump({"doc": "primes q in [2^31, 2^48) with q | T(n) (T(0)=T(1)=0, T(2)=1)",
                   "pairs": [[n, q] for n, q in pairs]}, f, indent=1)
        f.write("\n")
    print(f"{len(pairs)} pairs -> {out}")


if __name__ == "__main__":
    main()
