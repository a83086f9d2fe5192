"""Writes trib_near_residues.json: primes 2^36 <= q < 2^44 with T(n) = r (mod q)
for a small nonzero r, i.e. q | T(n) - r, for 60 <= n < 400 and r in
{1, -1, 3, -7} (T(0) = T(1) = 0, T(2) = 1), found with sympy's elliptic-curve
factoring (this container only).

Each triple (n, r, q) is a fact about Tribonacci numbers, checked here with
Python integers ((T(n) - r) % q == 0).  The Tribonacci lazy form tests most
of its unreduced terms x with a one-FMA filter that passes every multiple of
q and also the x within about q 2^-33 of one; an index whose term is r mod q
for a small r != 0 at q > |r| 2^33 is such a near miss, which the exact
re-test must reject.  The GPU tests place these indices at every position of
the blocks.
"""
import json
import os
import time

import sympy
from sympy.ntheory import ecm, isprime


def main(budget_s=300.0):
    t0 = time.time()
    T = [0, 0, 1]
    while len(T) < 400:
        T.append(T[-1] + T[-2] + T[-3])
    found = {}
    for n in range(60, 400):
        for r in (1, -1, 3, -7):
            if time.time() - t0 > budget_s:
                break
            v = abs(T[n] - r)
            for q in sympy.primerange(2, 20000):
                while v % q == 0:
                    v //= q
            if v == 1:
                continue
            try:
                fs = {v} if isprime(v) else ecm(v, B1=20000, B2=2000000, max_curve=40, seed=n)
            except ValueError:  # no split within these bounds: skip
                continue
            for q in fs:
                q = int(q)
                if isprime(q) and (1 << 36) <= q < (1 << 44) and q > abs(r) << 34:
                    found.setdefault(q.bit_length(), []).append((n, r, q))
    triples = sorted({t for ts in found.values() for t in ts[:3]}, key=lambda z: z[2])
    for n, r, q in triples:
        assert (T[n] - r) % q == 0 and T[n] % q != 0
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trib_near_residues.json")
    with open(out, "w") as f:
        json.dump({"doc": "primes q in [2^36, 2^44) with T(n) = r (mod q) for a small r != 0 "
                          "(T(0)=T(1)=0, T(2)=1); at most three per bit length",
                   "triples": [[n, r, q] for n, r, q in triples]}, f, indent=1)
        f.write("\n")
    print(f"{len(triples)} triples -> {out}")


if __name__ == "__main__":
    main()
