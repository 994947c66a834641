"""Finds prime factors p of n! + c and n!/n# + c (c = +-1, 101 <= n < 200;
n composite for n!/n#) with 2^34 < p < 2^52: moduli that the walk takes in its
narrow form (plain residues below 2p, pfc::narrow_mul) for multipliers below
300.  Each number is searched by Brent's variant of Pollard rho (batched gcd)
for a bounded number of iterations, so the output is a subset of those
factors.  The vectors are written to narrow_factors.json; the search takes a
few minutes on 8 processes and is not run by the tests, which only check and
use its output.

    python tests/golden/find_narrow_factors.py
"""
import json
import math
import multiprocessing
import os

import sympy
from sympy import primerange

LOW, HIGH = 2**34, 2**52
NMIN, NMAX = 101, 200
ITERS = 1 << 18
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "narrow_factors.json")


def brent_rho(n, c, iters):
    """A nontrivial factor of the composite n found within `iters` steps of
    x <- x^2 + c, or None."""
    y, r, q, g, x, ys = 2, 1, 1, 1, 2, 2
    done = 0
    while g == 1 and done < iters:
        x = y
        for _ in range(r):
            y = (y * y + c) % n
        k = 0
        while k < r and g == 1:
            ys = y
            for _ in range(min(128, r - k)):
                y = (y * y + c) % n
                q = q * abs(x - y) % n
            g = math.gcd(q, n)
            k += 128
        done += r
        r *= 2
    if g == n:
        g = 1
        while g == 1:
            ys = (ys * ys + c) % n
            g = math.gcd(abs(x - ys), n)
    return g if 1 < g < n else None


def factors_between(y, low, high):
    """Prime factors of y in (low, high) that rho finds within its budget."""
    out, stack = [], [y]
    while stack:
        m = stack.pop()
        if m < 2:
            continue
        if sympy.isprime(m):
            if low < m < high:
                out.append(int(m))
            continue
        g = None
        for c in (1, 3):
            g = brent_rho(m, c, ITERS)
            if g:
                break
        if g:
            stack += [g, m // g]
    return sorted(set(out))


def search(job):
    kind, n, c = job
    f = math.factorial(n)
    if kind == "C":
        f //= math.prod(primerange(2, n + 1))
    return [{"kind": kind, "n": n, "c": c, "p": p} for p in factors_between(f + c, LOW, HIGH)]


def main():
    jobs = []
    for n in range(NMIN, NMAX):
        for c in (-1, 1):
            jobs.append(("F", n, c))
            if not sympy.isprime(n):
                jobs.append(("C", n, c))
    with multiprocessing.Pool(8) as pool:
        vecs = [v for part in pool.map(search, jobs, chunksize=2) for v in part]
    for v in vecs:
        print(v)
    with open(OUT, "w") as f:
        json.dump({"nmin": NMIN, "nmax": NMAX, "note": "found by find_narrow_factors.py (a subset of the "
                   "prime factors in (2^34, 2^52))", "vectors": vecs}, f, indent=1)


if __name__ == "__main__":
    main()
