"""Finds prime factors p of n# + c (c = +-1, n a prime in [101, 400)) above
2^14 * 397^2, where a primorial walk over the primes in [101, 400) tests its
steps in triples.  The vectors are written to primorial_factors.json; the
search (sympy's factorint up to 10^6, then Pollard rho) takes a few minutes
and is not run by the tests, which only check and use its output.

    python tests/golden/find_primorial_factors.py
"""
import json
import os

import sympy
from sympy import primerange, primorial
from sympy.ntheory import pollard_rho

LIMIT = 2**14 * 397**2          # the triple-step bound for multipliers up to 397
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "primorial_factors.json")


def factors_above(y, low):
    """Prime factors of y above low that trial division to 10^6 and a few
    Pollard-rho runs find (not a complete factorisation)."""
    rest = y
    for q, e in sympy.factorint(y, limit=10**6).items():
        if q < 10**6:
            rest //= q**e
    out = []
    for seed in range(3):
        if rest == 1 or sympy.isprime(rest):
            break
        d = pollard_rho(rest, seed=seed, max_steps=300000)
        if d and 1 < d < rest:
            for g in (d, rest // d):
                if sympy.isprime(g) and low < g < 2**64:
                    out.append(int(g))
            rest //= d
    return out


def main():
    primes = list(primerange(101, 400))
    vecs = []
    for i, n in enumerate(primes):
        x = primorial(n, nth=False)
        for c in (-1, 1):
            for p in factors_above(x + c, LIMIT):
                vecs.append({"n": int(n), "index": i, "c": c, "p": p})
                print(vecs[-1], flush=True)
    with open(OUT, "w") as f:
        json.dump({"nmin": 101, "nmax": 400, "vectors": vecs}, f, indent=1)


if __name__ == "__main__":
    main()
