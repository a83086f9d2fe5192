"""Writes fib_large_factors.json: primes 2^37 <= q < 2^47 dividing F(m) for
100 <= m < 1500, found with sympy's Pollard rho (this container only).

Each pair (m, q) is a fact about Fibonacci numbers, checked here with Python
integers (F(m) % q == 0).  The GPU tests use them to plant hits at large
primes, where a random prime window holds none.
"""
import json
import os
import time

import sympy
from sympy.ntheory import isprime, pollard_rho


def main(budget_s=240.0):
    t0 = time.time()
    F = [0, 1]
    for _ in range(2, 1500):
        F.append(F[-1] + F[-2])
    found = set()
    for m in range(100, 1500):
        if time.time() - t0 > budget_s:
            break
        n = F[m]
        for q in sympy.primerange(2, 5000):
            while n % q == 0:
                n //= q
        tries = 0
        while n > 1 and not isprime(n) and tries < 3:
            d = pollard_rho(n, retries=1, max_steps=1 << 16, seed=tries + m)
            tries += 1
            if not d:
                continue
            for x in (d, n // d):
                if isprime(x) and (1 << 37) <= x < (1 << 47):
                    found.add((m, x))
            n = n // d if isprime(d) else d
        if isprime(n) and (1 << 37) <= n < (1 << 47):
            found.add((m, n))
    pairs = sorted(found, key=lambda z: z[1])
    for m, q in pairs:
        assert F[m] % q == 0
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fib_large_factors.json")
    with open(out, "w") as f:
        json.dump({"doc": "primes q in [2^37, 2^47) with q | F(m) (F(0)=0, F(1)=1)",
                   "pairs": [[m, q] for m, q in pairs]}, f, indent=1)
        f.write("\n")
    print(f"{len(pairs)} pairs -> {out}")


if __name__ == "__main__":
    main()
