"""Seeded job mixes for the batched and served forms.

A unit (job) is drawn as:
  * mode in {-!, -#, -c} with equal odds;
  * nmin = 101 + floor(U[0, 1e6]), n-width = floor(exp(U[ln 1e3, ln 1e6]));
  * pmin = floor(exp(U[ln 1e10, ln 2^63])), p-width = floor(exp(U[ln 1e6, ln 1e9])).
The generator is splitmix64 so that C++ and Python agree on the draws.

A primorial unit whose nmax is prime would stop the reference with "CPU sieve
failure" (cl_sieve.cpp:1058-1066); such units get nmax moved up to the next
composite and are marked ``nmax_adjusted``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))


def _is_prime_small(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    r = int(math.isqrt(n))
    f = 3
    while f <= r:
        if n % f == 0:
            return False
        f += 2
    return True


@dataclass
class Unit:
    modes: int
    pmin: int
    pmax: int
    nmin: int
    nmax: int
    nmax_adjusted: bool = False

    def desc(self):
        return (self.modes, self.pmin, self.pmax, self.nmin, self.nmax)


def draw_unit(rng: SplitMix64) -> Unit:
    modes = (1, 2, 4)[min(2, int(rng.uniform() * 3))]
    nmin = 101 + int(rng.uniform() * 1e6)
    nw = int(math.exp(math.log(1e3) + rng.uniform() * (math.log(1e6) - math.log(1e3))))
    pmin = int(math.exp(math.log(1e10) + rng.uniform() * (math.log(2.0 ** 63) - math.log(1e10))))
    pw = int(math.exp(math.log(1e6) + rng.uniform() * (math.log(1e9) - math.log(1e6))))
    u = Unit(modes, pmin, pmin + pw, nmin, nmin + nw)
    if modes == 2:
        while _is_prime_small(u.nmax):
            u.nmax += 1
            u.nmax_adjusted = True
    return u


def mix(n: int, seed: int = 12345) -> List[Unit]:
    rng = SplitMix64(seed)
    return [draw_unit(rng) for _ in range(n)]


def walk_steps(u: Unit) -> int:
    """Multiplier steps per candidate: n-range for F, composites for C, primes for P."""
    lo, hi = u.nmin, u.nmax
    sieve = bytearray([1]) * (hi + 1)
    sieve[0] = sieve[1] = 0
    for i in range(2, int(math.isqrt(hi)) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(range(i * i, hi + 1, i)))
    if u.modes == 2:
        return sum(sieve[lo:hi])            # primes in [nmin, nmax-1]
    if u.modes == 4:
        return (hi - lo) - sum(sieve[lo:hi])
    return hi - lo


def poisson_arrivals(n: int, rate: float, seed: int) -> List[float]:
    """Arrival times (seconds) of a Poisson process with the given rate."""
    rng = SplitMix64(seed ^ 0x5DEECE66D)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(1.0 - rng.uniform()) / rate
        out.append(t)
    return out
