"""Pins the CPU oracle (oracle/pfc_oracle.c) before it is trusted as the
checker: against the reference's own golden vectors and constants, against
the reference's verifyprime.cpp / putil.c built from source into oracle/_ref,
and against independent Python big-integer mathematics on small inputs."""
import math
import random

import numpy as np
import pytest
import sympy

F, P, Cm = 1, 2, 4
R = 1 << 64
SIEVE = [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97,
         101, 103, 107, 109, 113]


# ----------------------------------------------------------------- Montgomery


def test_montgomery_kat_2_64_minus_59(orc):
    """verify.cl:46 holds p = 2^64-59 with q, r2, one, two, nmo."""
    p = 18446744073709551557
    assert orc.invert(p) == 3751880150584993549
    one = (R % p)
    assert one == 59
    rec, _ = orc.candidate_record(F, p, 101, 102)
    assert rec[1] == 3751880150584993549 and rec[2] == 3481 and rec[3] == 59
    assert rec[5] == 18446744073709551498


def test_mmul_brute_force(orc):
    rng = random.Random(5)
    for i in range(3000):
        p = rng.choice([rng.randrange(3, R) | 1, R - rng.randrange(1, 1000) | 1, rng.randrange(3, 1 << 20) | 1])
        if p >= R:
            p -= 2
        q = orc.invert(p)
        assert (p * q) % R == 1
        a, b = rng.randrange(p), rng.randrange(p)
        assert orc.mmul(a, b, p, q) == a * b * pow(R, -1, p) % p
        assert orc.madd(a, b, p) == (a + b) % p


def test_mmul_matches_reference_build(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = random.Random(6)
    for _ in range(2000):
        p = rng.randrange(3, R) | 1
        q = orc.ref_invert(p)
        assert q == orc.invert(p)
        a, b = rng.randrange(p), rng.randrange(p)
        assert orc.mmul(a, b, p, q) == orc.ref_mmul(a, b, p, q)


# ----------------------------------------------------------------- primality


def spsp2_bigint(n):
    t = ((n - 1) & -(n - 1)).bit_length() - 1
    d = (n - 1) >> t
    a = pow(2, d, n)
    if a in (1, n - 1):
        return True
    for _ in range(1, t):
        a = a * a % n
        if a == n - 1:
            return True
    return False


def test_strong_prp2_brute_force(orc):
    rng = random.Random(7)
    xs = list(range(3, 20000, 2)) + [rng.randrange(3, R) | 1 for _ in range(3000)]
    xs += [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581,
           85489, 88357, 90751, 3215031751, 2152302898747, 3474749660383]
    for x in xs:
        assert orc.strong_prp2(x) == spsp2_bigint(x), x


def test_fermat_form_shift_edge(orc):
    """N = 2^t + 1 makes the odd part of N-1 equal 1; the reference's
    masked shift and the exact test agree on every such N that reaches it."""
    for t in range(7, 64):
        N = (1 << t) + 1
        if N % 3 == 0 or N % 5 == 0 or any(N % q == 0 for q in SIEVE):
            continue
        assert orc.strong_prp2(N) == spsp2_bigint(N)


def test_is_prime_against_sympy_and_reference(orc):
    rng = random.Random(8)
    xs = list(range(3, 5000, 2)) + [rng.randrange(3, R) | 1 for _ in range(3000)]
    xs += [3215031751, 2152302898747, 3474749660383, 341550071728321, 3825123056546413051,
           18446744073709551557, (1 << 61) - 1]
    ref = orc.ref_available()
    for x in xs:
        e = sympy.isprime(x)
        assert orc.is_prime(x) == e, x
        if ref:
            assert orc.ref_is_prime(x) == e, x


# ----------------------------------------------------------------- prime lists, wheel


def test_primes_against_sympy(orc):
    for lo, hi in [(2, 1000), (101, 5000), (999000, 1001000), (2, 2), (4, 4), (10**7, 10**7 + 5000)]:
        assert list(orc.primes(lo, hi)) == list(sympy.primerange(lo, hi + 1))


def test_wheel_offset_properties(orc):
    gaps = [4, 2, 4, 2, 4, 6, 2, 6]
    rng = random.Random(9)
    for s in [114, 115, 119, 120, 121] + [rng.randrange(114, 1 << 62) for _ in range(500)]:
        N, k = orc.wheel_offset(s)
        assert N >= s and math.gcd(N, 30) == 1
        assert all(math.gcd(x, 30) != 1 for x in range(s, N))
        x = N
        for j in range(16):  # walking the wheel from (N, k) visits exactly the units mod 30
            nxt = x + gaps[(k + j) % 8]
            assert all(math.gcd(y, 30) != 1 for y in range(x + 1, nxt)) and math.gcd(nxt, 30) == 1
            x = nxt


def candidates_bigint(lo, hi):
    out = [p for p in sympy.primerange(max(lo, 3), min(hi, 114))]
    for c in range(max(lo, 114) | 1, hi, 2):
        if c % 3 and c % 5 and all(c % q for q in SIEVE) and spsp2_bigint(c):
            out.append(c)
    return out


@pytest.mark.parametrize("lo,hi", [(3, 20000), (101, 114), (114, 200), (10**9, 10**9 + 30000),
                                   ((1 << 62) - 20000, 1 << 62), (R - 20000, R - 1)])
def test_candidates_bruteforce(orc, lo, hi):
    assert list(orc.candidates(lo, hi)) == candidates_bigint(lo, hi)


def test_candidates_keep_pseudoprimes(orc):
    """spsp(2) with least prime factor > 113 are candidates; others are not."""
    c = set(int(x) for x in orc.candidates(3, 200000))
    for n in [42799, 49141, 88357, 104653, 130561, 196093]:   # spsp(2), lpf > 113
        assert spsp2_bigint(n) and min(sympy.factorint(n)) > 113 and n in c
    for n in [2047, 3277, 4033, 4681, 8321, 15841]:           # spsp(2), lpf <= 113
        assert n not in c


def test_golden_primecounts(orc, golden):
    """primecount of all 16 reference self-test ranges (cl_sieve.cpp:1787-2080)."""
    for g in golden:
        n = orc.lib().orc_candidates(g["pmin"], g["pmax"], None, 0)
        assert n == g["primecount"], g["case"]


# ----------------------------------------------------------------- tables


def test_power_table_product(orc):
    for N in [100, 1000, 4999, 25000]:
        prod, e, bits = orc.power_table(N)
        assert int(prod[0]) == 2
        x = 1
        for pr, ee, b in zip(prod, e, bits):
            x *= int(pr) ** int(ee)
            assert b == (0x80000000 if ee <= 1 else 1 << (int(ee).bit_length() - 2))
        assert x == math.factorial(N)
        assert all(int(v) < R for v in prod)


def test_prime_and_composite_products(orc):
    for N in [100, 1000, 9999]:
        pp = orc.prime_products(N)
        assert math.prod(int(x) for x in pp) == sympy.primorial(N, nth=False)
        cp = orc.comp_products(N)
        assert math.prod(int(x) for x in cp) == math.factorial(N) // sympy.primorial(N, nth=False)


def test_reference_kat_constants():
    """verifyprime.cpp:155,172,191 seed the slow check with 34!, 101#, 44!/44#."""
    f34 = (0xde1bc4d19efcac82 << 64) | 0x445da75b00000000
    p101 = (0xaf2fa8f8a2d02a93 << 64) | 0xae69c9f8987d5efe
    c44 = (0x98dcc10f185c0e67 << 64) | 0x3c93ff0000000000
    assert f34 == math.factorial(34)
    assert p101 == sympy.primorial(101, nth=False)
    assert c44 == math.factorial(44) // sympy.primorial(44, nth=False)


# ----------------------------------------------------------------- walk + checksum


def X(mode, n):
    if mode == F:
        return math.factorial(n)
    if mode == P:
        return sympy.primorial(n, nth=False) if n >= 2 else 1
    return math.factorial(n) // (sympy.primorial(n, nth=False) if n >= 2 else 1)


def closed_form(modes, pmin, pmax, nmin, nmax):
    """The job's outputs from first principles (big integers, sympy)."""
    cands = candidates_bigint(pmin, pmax)
    last = nmax - 1
    dev = 0
    facs = []
    walks = [m for m in (F, Cm) if modes & m] if not modes & P else [P]
    for c in cands:
        mont = lambda x: (x % c) * R % c
        if modes & P:
            term = mont(X(P, last))
        else:
            term = sum(mont(X(m, last)) for m in walks) + mont(last)
        dev += term
        for m in walks:
            for n in range(nmin, nmax):
                if m == P and not sympy.isprime(n):
                    continue
                if m == Cm and sympy.isprime(n):
                    continue
                r = X(m, n) % c
                if r == 1 or r == c - 1:
                    facs.append((c, -n if r == 1 else n, {F: 0, P: 1, Cm: 2}[m]))
    facs.sort(key=lambda f: (f[0], abs(f[1]), f[2]))
    cs = dev
    kept = []
    for f in facs:
        if sympy.isprime(f[0]):
            kept.append(f)
            cs += abs(f[1]) + (1 if f[1] > 0 else -1)
    return cs % R, len(cands), kept


@pytest.mark.parametrize("job", [(F, 101, 1500, 101, 700), (P, 101, 1500, 101, 700),
                                 (Cm, 101, 1500, 101, 700), (F | Cm, 101, 1500, 101, 700),
                                 (F, 10**6, 10**6 + 3000, 150, 400), (Cm, 10**9, 10**9 + 2000, 300, 500),
                                 (P, 3, 900, 101, 1200), (F | Cm, R - 3000, R - 1, 101, 160)])
def test_job_closed_form(orc, job):
    cs, pc, kept = closed_form(*job)
    r = orc.run(*job)
    assert r.primecount == pc
    assert r.checksum == cs
    assert [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors] == kept


def test_wilson_factors(orc):
    """Every prime p with p-1 in [nmin, nmax) divides (p-1)!+1 and p-2 gives (p-2)!-1."""
    r = orc.run(F, 101, 3000, 101, 3000)
    got = set((int(f["p"]), int(f["nc"])) for f in r.factors)
    for p in sympy.primerange(103, 3000):
        assert (p, p - 1) in got and (p, -(p - 2)) in got


def test_prime_values_have_no_factors(orc):
    """n!+1 is prime for n = 116, 154; n!-1 for n = 166: no p may divide them."""
    r = orc.run(F, 10**6, 10**6 + 20000, 101, 170)
    for f in r.factors:
        assert (int(f["nc"])) not in (116, 154, -166)


@pytest.mark.parametrize("case", [3, 10])
def test_golden_selftest_cpu_cases(orc, golden, case):
    """Self-test cases small enough for the CPU (cl_sieve.cpp:1831, 1959)."""
    g = golden[case - 1]
    r = orc.run(g["modes"], g["pmin"], g["pmax"], g["nmin"], g["nmax"])
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (g["factorcount"], g["primecount"], g["checksum"])


@pytest.mark.slow
@pytest.mark.parametrize("case", [6, 14])
def test_golden_selftest_cpu_cases_slow(orc, golden, case):
    g = golden[case - 1]
    r = orc.run(g["modes"], g["pmin"], g["pmax"], g["nmin"], g["nmax"])
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (g["factorcount"], g["primecount"], g["checksum"])


# ----------------------------------------------------------------- verify


def test_verify_against_reference_build(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    r = orc.run(F | Cm, 101, 3000, 101, 2000)
    for f in r.factors[:300]:
        p, nc, t = int(f["p"]), int(f["nc"]), int(f["type"])
        n, c = abs(nc), (-1 if nc < 0 else 1)
        assert orc.verify(p, n, c, t)
        assert orc.ref_verify(p, n, c, t, 2000)
        assert not orc.verify(p, n, -c, t)
        assert not orc.ref_verify(p, n, -c, t, 2000)
    r = orc.run(P, 101, 3000, 101, 2001)
    for f in r.factors[:200]:
        p, nc = int(f["p"]), int(f["nc"])
        assert orc.verify(p, abs(nc), -1 if nc < 0 else 1, 1)
        assert orc.ref_verify(p, abs(nc), -1 if nc < 0 else 1, 1, 2001)


def _primorial_fixtures():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "primorial_factors.json")
    with open(path) as f:
        return json.load(f)


def test_primorial_factor_fixtures(orc):
    """The planted primorial factors (tests/golden/primorial_factors.json):
    each p is prime, divides n# + c by big integers, lies above the bound
    2^14 * 397^2 where the device walk tests its steps in triples, and the
    oracle reports it in a small p-window around it."""
    d = _primorial_fixtures()
    assert (d["nmin"], d["nmax"]) == (101, 400)
    primes = [int(q) for q in sympy.primerange(101, 400)]
    assert {v["index"] % 3 for v in d["vectors"]} == {0, 1, 2}
    for v in d["vectors"]:
        n, c, p = v["n"], v["c"], v["p"]
        assert primes[v["index"]] == n
        assert sympy.isprime(p) and p > 2**14 * 397**2
        assert (sympy.primorial(n, nth=False) + c) % p == 0
        r = orc.run(2, p - 200, p + 200, 101, 400)
        got = [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]
        assert (p, c * n, 1) in got, v


def test_narrow_factor_fixtures(orc):
    """The planted factors of the narrow walk (tests/golden/narrow_factors.json):
    each p is prime, lies in (2^34, 2^52), divides n! + c (kind F) or
    n!/n# + c (kind C, n composite) by big integers, and the oracle reports
    it in a small p-window around it; the fixtures have hits of both signs
    at odd and even n in both chains."""
    import json
    import math
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "narrow_factors.json")
    with open(path) as f:
        d = json.load(f)
    assert (d["nmin"], d["nmax"]) == (101, 200)
    seen = set()
    for v in d["vectors"]:
        kind, n, c, p = v["kind"], v["n"], v["c"], v["p"]
        assert sympy.isprime(p) and 2**34 < p < 2**52
        x = math.factorial(n)
        if kind == "C":
            assert not sympy.isprime(n)
            x //= sympy.primorial(n, nth=False)
        assert (x + c) % p == 0, v
        r = orc.run(1 if kind == "F" else 4, p - 200, p + 200, 101, 300)
        got = [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]
        assert (p, c * n, 0 if kind == "F" else 2) in got, v
        seen.add((kind, n % 2, c))
    assert {(k, par) for k, par, _ in seen} == {("F", 0), ("F", 1), ("C", 0), ("C", 1)}
    assert {c for _, _, c in seen} == {-1, 1}
