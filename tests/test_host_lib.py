"""CPU-side checks of libpfcsieve.so: the exported C ABI, its host functions
(isPrime, number grammar, harvest, factor lines, setup tables) and the device
arithmetic compiled for the host.  No GPU calls."""
import ctypes as C
import math
import os
import random
import re
import subprocess

import numpy as np
import pytest
import sympy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
R = 1 << 64


@pytest.fixture(scope="module")
def L():
    from pfcsieve_amd._native import lib
    return lib()


def test_every_header_symbol_exported(L):
    hdr = open(os.path.join(ROOT, "include", "pfc.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(pfc_\w+)\(", hdr, re.M))
    assert len(names) >= 25
    from pfcsieve_amd._native import SIGNATURES
    assert names == set(SIGNATURES), names ^ set(SIGNATURES)
    for n in names:
        assert hasattr(L, n), n


def test_is_prime(L):
    rng = random.Random(1)
    xs = list(range(0, 3000)) + [rng.randrange(3, R) for _ in range(3000)]
    xs += [3215031751, 3825123056546413051, 18446744073709551557, 2**61 - 1, 2**64 - 1]
    for x in xs:
        assert bool(L.pfc_is_prime(x)) == sympy.isprime(x), x


GRAMMAR = ["101", "1e6", "1E6", "25M", "2G", "3T", "1P", "5K", "1k", "1m", "2g", "1t", "1p", "1b10",
           "3B4", "0x10", "010", "12abc", "1e", "1ee", "1KK", "18446744073709551615", "18446744073709551616",
           "2e19", "1e20", "17b60", "1b64", "", "-5", "4611686018427387904", "100e7", "1730720716e6"]


def test_parse_grammar_against_reference(L, orc):
    for s in GRAMMAR:
        for lo, hi in [(3, R - 2), (4, R - 1), (101, 0x7FFFFFFE), (1, 128)]:
            out = C.c_uint64()
            rc = L.pfc_parse_uint64(s.encode(), lo, hi, C.byref(out))
            if orc.ref_available():
                rrc, rv = orc.ref_parse_uint64(s, lo, hi)
                assert rc == rrc, (s, lo, hi)
                if rc == 0:
                    assert out.value == rv, s


def test_parse_grammar_values(L):
    def p(s, lo=0, hi=R - 1):
        out = C.c_uint64()
        rc = L.pfc_parse_uint64(s.encode(), lo, hi, C.byref(out))
        return out.value if rc == 0 else rc
    assert p("1e6") == 10**6 and p("25M") == 25 * 10**6 and p("3T") == 3 * 10**12
    assert p("1k") == 1024 and p("1b10") == 1024 and p("0x10") == 16 and p("010") == 8
    assert p("12abc") == -1 and p("1e20") == -2 and p("50", lo=101) == -2


def test_format_factor(L):
    from pfcsieve_amd._native import Factor
    buf = C.create_string_buffer(96)
    for f, line in [((1000003, -200, 0), "1000003 | 200!-1\n"), ((17, 5, 1), "17 | 5#+1\n"),
                    ((2**64 - 59, 300, 2), "18446744073709551557 | 300!/#+1\n")]:
        L.pfc_format_factor(C.byref(Factor(*f)), buf, 96)
        assert buf.value.decode() == line


def test_harvest_matches_oracle(L, orc):
    from pfcsieve_amd._native import Factor
    # raw records of a job with pseudoprime-free small p, shuffled
    r = orc.run(1 | 4, 101, 5000, 101, 3000)
    raw = [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]
    # add records of a base-2 pseudoprime (composite p) that must be dropped
    raw += [(42799, 1234, 0), (42799, -2000, 2)]
    random.Random(2).shuffle(raw)
    arr = (Factor * len(raw))(*[Factor(*x) for x in raw])
    cs = C.c_uint64(r.dev_checksum)
    kept = C.c_size_t()
    assert L.pfc_harvest(arr, len(raw), 0, C.byref(cs), C.byref(kept)) == 0
    assert cs.value == r.checksum
    assert [(arr[i].p, arr[i].nc, arr[i].type) for i in range(kept.value)] == \
        [(int(f["p"]), int(f["nc"]), int(f["type"])) for f in r.factors]


def test_harvest_cpu_verify_rejects_bad_record(L):
    from pfcsieve_amd._native import Factor
    good = [Factor(103, 102, 0), Factor(103, -101, 0)]          # Wilson: 102!+1, 101!-1
    arr = (Factor * 2)(*good)
    cs, kept = C.c_uint64(0), C.c_size_t()
    assert L.pfc_harvest(arr, 2, 2, C.byref(cs), C.byref(kept)) == 0 and kept.value == 2
    bad = (Factor * 1)(Factor(103, 101, 0))
    assert L.pfc_harvest(bad, 1, 2, C.byref(cs), C.byref(kept)) == -6
    assert b"is not a factor of 101!+1" in L.pfc_last_error()


def test_verify_factor(L):
    assert L.pfc_verify_factor(103, 102, 1, 0) and L.pfc_verify_factor(103, 101, -1, 0)
    assert not L.pfc_verify_factor(103, 101, 1, 0)
    for n in range(101, 400):
        for p in [1000003, 10007]:
            for t, X in [(0, math.factorial(n)), (1, sympy.primorial(n, nth=False)),
                         (2, math.factorial(n) // sympy.primorial(n, nth=False))]:
                r = X % p
                assert bool(L.pfc_verify_factor(p, n, -1, t)) == (r == 1)
                assert bool(L.pfc_verify_factor(p, n, 1, t)) == (r == p - 1)


@pytest.mark.parametrize("kind,N", [(0, 100), (0, 4999), (0, 100000), (1, 100), (1, 99999),
                                    (2, 100), (2, 77777)])
def test_setup_table_reconstructs_start_value(L, kind, N):
    """The table's documented evaluation (Horner over base-2^k digit
    positions, suffix products over digits) gives X(N) exactly."""
    from pfcsieve_amd.sieve import TableLayout
    t = TableLayout(kind, N)
    prim = sympy.primorial(N, nth=False)
    expect = [math.factorial(N), prim, math.factorial(N) // prim][kind]
    assert t.value() == expect
    assert all(1 < c < 2**63 for c in t.chunks)
    assert 1 <= t.digit_bits <= 8 and len(t.groups) == t.npos * ((1 << t.digit_bits) - 1) + 1


@pytest.mark.parametrize("kind,N", [(0, 3000), (0, 20000), (2, 20000), (1, 5000)])
def test_setup_table_exponents_are_legendre(L, kind, N):
    """Factoring every chunk back into primes: a prime r in group (b, t)
    contributes t * B^b to its exponent, and the sum over groups is its
    exponent in X(N) (Legendre's formula; minus one for N!/N#; one for N#)."""
    from pfcsieve_amd.sieve import TableLayout
    t = TableLayout(kind, N)
    B = 1 << t.digit_bits
    got = {}
    for b in range(t.npos):
        for d in range(1, B):
            for c in t.group(b, d):
                for r, mult in sympy.factorint(c).items():
                    assert mult == 1 and r <= N               # distinct primes within a chunk
                    got[r] = got.get(r, 0) + d * B ** b
    want = {}
    for r in sympy.primerange(2, N + 1):
        e = 1 if kind == 1 else sum(N // r ** i for i in range(1, 40) if r ** i <= N) - (kind == 2)
        if e:
            want[r] = e
    assert got == want


@pytest.mark.parametrize("kind,N,ratio", [(0, 999999, 0.8), (2, 999999, 0.75), (0, 99999, 0.9)])
def test_setup_table_digit_width_saves_products(L, kind, N, ratio):
    """At large N the chosen digit width needs fewer modular products than
    binary digits (each prime multiplied in popcount(e) times): 999999! from
    34,892 to about 26,200."""
    from pfcsieve_amd.sieve import TableLayout
    t = TableLayout(kind, N)
    assert t.digit_bits > 1
    # the binary layout's cost: every prime once per set exponent bit
    pr = list(sympy.primerange(2, N + 1))
    binary_bits = 0
    for r in pr:
        e = sum(N // r ** i for i in range(1, 40) if r ** i <= N) - (kind == 2)
        binary_bits += bin(e).count("1") * math.log2(r)
    assert t.products() < ratio * binary_bits / 60


# ----------------------------------------------------------------- device arithmetic on the host


@pytest.fixture(scope="module")
def A():
    out = os.path.join(ROOT, "tests", "host", "_build", "libarith.so")
    src = os.path.join(ROOT, "tests", "host", "arith_shim.cpp")
    hdrs = [os.path.join(ROOT, "pfcsieve_amd", "csrc", h) for h in ("pfc_arith.h", "pfc_share.h")]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(f) for f in [src] + hdrs):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        # clang, as on the device (the same builtins are compiled)
        cxx = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") else "g++"
        subprocess.run([cxx, "-O2", "-fPIC", "-shared", "-o", out, src], check=True)
    L = C.CDLL(out)
    u64, u32 = C.c_uint64, C.c_uint32
    L.t_modulus.argtypes = [u64, C.POINTER(u32 * 4)]
    for n, args in [("t_mulmod32", [u64, u64, u32]), ("t_mulmod64", [u64, u64, u64]),
                    ("t_sqrmod", [u64, u64]), ("t_mont_one", [u64])]:
        getattr(L, n).argtypes = args
        getattr(L, n).restype = u64
    L.t_modulus_ok.argtypes = [u64]
    L.t_strong_prp2.argtypes = [u64]
    L.t_strong_prp2_pair.argtypes = [u64, u64, C.POINTER(C.c_int * 2)]
    L.t_is_prime_u32.argtypes = [u32]
    L.t_generate.argtypes = [u64, u64, C.POINTER(u64), C.c_size_t]
    L.t_generate.restype = C.c_size_t
    L.t_start_residue.argtypes = [C.POINTER(u64), C.POINTER(u32), u32, u32, u64]
    L.t_start_residue.restype = u64
    L.t_start_residue_split.argtypes = [C.POINTER(u64), C.POINTER(u32), u32, u32, u64, u32]
    L.t_start_residue_split.restype = u64
    L.t_to_mont.argtypes = [u64, u64]
    L.t_to_mont.restype = u64
    L.t_walk_step.argtypes = [u64, u64, u32, C.POINTER(u64), C.POINTER(C.c_int)]
    L.t_walk_step.restype = u64
    L.t_walk_attention.argtypes = [u64, u64, u64]
    L.t_mulmod64_nofix.argtypes = [u64, u64, u64, C.POINTER(C.c_int)]
    L.t_mulmod64_nofix.restype = u64
    L.t_narrow_consts.argtypes = [u64, C.POINTER(u32 * 3)]
    L.t_narrow_fits.argtypes = [u64, u32]
    L.t_narrow_walk.argtypes = [u64, u64, C.POINTER(u32), u32]
    L.t_narrow_walk.restype = u64
    L.t_narrow_attention.argtypes = [u64, u64, u64]
    L.t_medium_consts.argtypes = [u64, C.POINTER(u32 * 3)]
    L.t_medium_fits.argtypes = [u64, u32]
    L.t_medium_walk.argtypes = [u64, u64, C.POINTER(u32), u32]
    L.t_medium_walk.restype = u64
    L.t_share_steps.argtypes = [C.POINTER(u64), C.POINTER(u64), C.c_size_t, u64, u64, C.POINTER(u64)]
    return L


def _primes_mix(rng, k):
    out = [3, 5, 7, 113, 65537, 2**32 - 5, 2**32 + 15, 2**61 - 1, 2**62 - 57, 2**63 - 25, 2**63 + 29,
           2**64 - 59, 1000003, 10**12 + 39]
    out += [rng.randrange(3, R) | 1 for _ in range(k)]
    return [p if p < R else p - 2 for p in out]


def test_modulus_and_reciprocal(A):
    rng = random.Random(3)
    for p in _primes_mix(rng, 2000):
        m = (C.c_uint32 * 4)()
        A.t_modulus(p, C.byref(m))
        s = 64 - p.bit_length()
        d = p << s
        assert m[3] == s and (m[0] << 32 | m[1]) == d
        assert m[2] == ((1 << 96) - 1) // d - (1 << 32)
        assert A.t_modulus_ok(p)
        assert A.t_mont_one(p) == (R % p) << s


def test_mulmod_against_bigint(A):
    rng = random.Random(4)
    for p in _primes_mix(rng, 3000):
        s = 64 - p.bit_length()
        for a in [0, 1, p - 1, rng.randrange(p), rng.randrange(p)]:
            for n in [1, 2, 0x7FFFFFFF, 0xFFFFFFFF, rng.randrange(1 << 32)]:
                assert A.t_mulmod32(p, a << s, n) == (a * n % p) << s
            b = rng.randrange(R)
            assert A.t_mulmod64(p, a << s, b) == (a * b % p) << s
            assert A.t_sqrmod(p, a << s) == (a * a % p) << s
            assert A.t_to_mont(p, a << s) == (a * R % p) << s


def test_mulmod64_unreduced_input(A):
    """The start-residue chunk product: for b < 2^63 any 64-bit input is
    accepted, unreduced ones (x << s) + d included; the result is a * b mod p,
    shifted, or that plus d with the fix flag set."""
    rng = random.Random(13)
    fix = C.c_int()
    for p in _primes_mix(rng, 2000):
        s = 64 - p.bit_length()
        d = p << s
        for _ in range(4):
            x = rng.randrange(p)
            a = x << s
            if a + d < R and rng.random() < 0.5:
                a += d
            for b in [1, 2**63 - 1, rng.randrange(2**63), rng.randrange(2**32)]:
                got = A.t_mulmod64_nofix(p, a, b, C.byref(fix))
                want = (x * b % p) << s
                assert got in (want, want + d) and bool(fix.value) == (got >= d), (p, a, b)


def test_walk_attention_bounds(A):
    """The walk's high-word test against a band B passes every x <= B and
    every x >= d - B (which includes the unreduced x >= d), and passes an x
    strictly between only when its high word equals that of B or of d - B
    (then exactly those); checked at both edges of both intervals and at
    random x, for B = 2^s (+-1) and for paired and triple bands b << s, up to
    bands too wide for the high word, where every x passes."""
    rng = random.Random(11)
    for p in _primes_mix(rng, 500):
        s = 64 - p.bit_length()
        d, one = p << s, 1 << s
        bands = [one, rng.randrange(1, max(2, p // 2)) << s, (p // 2 - 1) << s if p > 4 else one]
        for B in bands:
            bh, dh = B >> 32, (d - B) >> 32
            edges = [0, 1, B - 1, B, B + 1, d - B - 1, d - B, d - B + 1, d - 1, d, d + 1, R - 1,
                     (bh << 32) | 0xFFFFFFFF, dh << 32]
            for x in edges + [rng.randrange(R) for _ in range(20)] + [rng.randrange(d) for _ in range(20)]:
                x %= R
                got = A.t_walk_attention(p, B, x)
                if x <= B or x >= d - B:
                    assert got == 1, (p, B, x)
                elif dh > bh:
                    assert got == int((x >> 32) in (bh, dh)), (p, B, x)
                else:
                    assert got == 1, (p, B, x)


def test_walk_step_against_bigint(A):
    """One walk step (3-by-2 step, attention test, correction) against big
    integers: the settled residue is a*n mod p; the raw one is that or that
    plus d; every raw value that is +-1 or unreduced raises attention; and
    steps landing exactly on +1 and -1 (n chosen as a quotient of residues)
    are caught."""
    rng = random.Random(12)
    raw, att = C.c_uint64(), C.c_int()
    # (p, a, n) whose raw result needs the second correction (found by a
    # random search over odd moduli and 32-bit n; it is far rarer for n below
    # 2^24 and random moduli).
    for p, a, n in [(11808775468092489729, 10992610456871730286, 4294967044),
                    (9248424854364649175, 5733249779935437510, 4294967082),
                    (9264490322874859521, 4812142961238367030, 3106929992),
                    (9275621306045530121, 5000254257389108843, 3738354044),
                    (9583256094208529071, 5956682225011036978, 3496521807)]:
        got = A.t_walk_step(p, a, n, C.byref(raw), C.byref(att))
        assert raw.value >= p and att.value and got == a * n % p, (p, a, n)
    for p in _primes_mix(rng, 1500):
        s = 64 - p.bit_length()
        d = p << s
        cases = [(rng.randrange(1, p), rng.randrange(1, min(p, 1 << 32))) for _ in range(30)]
        for target in (1, p - 1):
            for _ in range(4):
                n = rng.randrange(1, min(p, 1 << 32))
                if math.gcd(n, p) == 1:    # the moduli of the mix are odd, not all prime
                    cases.append((target * pow(n, -1, p) % p, n))
        for a, n in cases:
            got = A.t_walk_step(p, a << s, n, C.byref(raw), C.byref(att))
            want = a * n % p
            assert got == want << s, (p, a, n)
            assert raw.value in (want << s, (want << s) + d), (p, a, n)
            if raw.value >= d or want in (1, p - 1):
                assert att.value, (p, a, n)


def _narrow_moduli(rng, k):
    """Odd moduli across the narrow range (2^34, 2^55): both ends, near powers
    of two, and random bit lengths."""
    out = [2**34 + 1, 2**34 + 3, 2**35 - 1, 2**40 - 87, 10**12 + 39, 2**48 + 21, 2**55 - 55, 2**55 - 1,
           (2**41 // 3) | 1, 2**52 - 1]
    for _ in range(k):
        b = rng.randrange(35, 56)
        out.append(rng.randrange(2**(b - 1) + 1, 2**b) | 1)
    return [p for p in out if 2**34 < p < 2**55]


def _narrow_bmax(p):
    """The largest multiplier the narrow walk takes for p: p * n < 2^62, n < 2^22."""
    return min((2**62 - 1) // p, 2**22 - 1)


def test_narrow_fits_bounds(A):
    """narrow_fits holds exactly for 2^34 < p < 2^55, 1 <= bmax < 2^22 and
    p * bmax < 2^62."""
    rng = random.Random(21)
    for p in _narrow_moduli(rng, 300) + [2**34 - 1, 2**34, 2**55, 2**55 + 1, 2**60 + 1]:
        for b in [0, 1, 2, 2**22 - 1, 2**22, 2**32 - 1, (2**62 - 1) // p, (2**62 - 1) // p + 1,
                  rng.randrange(1, 2**22)]:
            b = min(b, 2**32 - 1)
            want = 2**34 < p < 2**55 and 1 <= b < 2**22 and p * b < 2**62
            assert A.t_narrow_fits(p, b) == int(want), (p, b)


def test_narrow_constants(A):
    """m = floor(2^64 / p) from the 3-by-2 reciprocal, np0 = 2^32 - p0 and
    kp1 = -(p1 + 1) mod 2^32 (a signed 24-bit value)."""
    rng = random.Random(22)
    out = (C.c_uint32 * 3)()
    for p in _narrow_moduli(rng, 2000):
        A.t_narrow_consts(p, C.byref(out))
        assert out[0] == R // p, p
        assert out[1] == (1 << 32) - (p & 0xFFFFFFFF), p
        assert out[2] == (-((p >> 32) + 1)) % (1 << 32), p
        assert (p >> 32) + 1 <= 1 << 23


def test_narrow_step_against_bigint(A):
    """The narrow step maps every X in [0, 2p) and n up to the largest the
    walk takes to a value in [0, 2p) congruent to X * n mod p; checked at the
    ends of both ranges (X = 0, 1, p - 1, p, 2p - 1; n = 1, 2, the maximum)
    and on chains of random multipliers, each step against big integers."""
    rng = random.Random(23)
    for p in _narrow_moduli(rng, 1500):
        bmax = _narrow_bmax(p)
        xs = [0, 1, 2, p - 1, p, p + 1, 2 * p - 2, 2 * p - 1] + [rng.randrange(2 * p) for _ in range(6)]
        ns = [1, 2, 3, bmax, bmax - 1, max(1, bmax // 2)] + [rng.randrange(1, bmax + 1) for _ in range(4)]
        for x in xs:
            for n in ns:
                got = A.t_narrow_walk(p, x, (C.c_uint32 * 1)(n), 1)
                assert 0 <= got < 2 * p and got % p == x * n % p, (p, x, n, got)
        chain = [rng.randrange(1, bmax + 1) for _ in range(64)]
        x = rng.randrange(2 * p)
        got = A.t_narrow_walk(p, x, (C.c_uint32 * 64)(*chain), 64)
        want = x
        for n in chain:
            want = want * n % p
        assert got < 2 * p and got % p == want, (p, x)


def test_narrow_step_near_multiples(A):
    """Products t = X * n just below and above multiples of p (where a wrong
    quotient estimate would show), including t close to 2^63."""
    rng = random.Random(24)
    for p in _narrow_moduli(rng, 800):
        bmax = _narrow_bmax(p)
        for _ in range(12):
            n = rng.choice([bmax, rng.randrange(1, bmax + 1)])
            k = rng.randrange(1, 2 * n)                  # a multiple k * p below 2p * n
            for delta in (-2, -1, 0, 1, 2):
                t = k * p + delta
                if t < 0 or t % n:
                    continue
                x = t // n
                if x >= 2 * p:
                    continue
                got = A.t_narrow_walk(p, x, (C.c_uint32 * 1)(n), 1)
                assert got < 2 * p and got % p == t % p, (p, x, n)
            # X chosen so that X * n mod p lands on 1 and p - 1 (hits), from both halves of [0, 2p)
            if math.gcd(n, p) == 1:
                for target in (1, p - 1):
                    x = target * pow(n, -1, p) % p
                    for xx in (x, x + p):
                        got = A.t_narrow_walk(p, xx, (C.c_uint32 * 1)(n), 1)
                        assert got < 2 * p and got % p == target, (p, xx, n)


def test_narrow_attention_bounds(A):
    """The narrow band test passes exactly the reduced x <= B and x >= p - B,
    for B = 1, pair bands and triple bands; checked at every interval edge
    and at random x."""
    rng = random.Random(25)
    for p in _narrow_moduli(rng, 400):
        bmax = _narrow_bmax(p)
        for B in [1, bmax, rng.randrange(1, bmax + 1), min(bmax * bmax, (p - 2) // 2)]:
            edges = [0, 1, B - 1, B, B + 1, p - B - 1, p - B, p - B + 1, p - 2, p - 1]
            for x in edges + [rng.randrange(p) for _ in range(30)]:
                if not 0 <= x < p:
                    continue
                assert A.t_narrow_attention(p, B, x) == int(x <= B or x >= p - B), (p, B, x)


def _medium_moduli(rng, k):
    """Odd moduli across the medium range (2^32, 2^63): both ends, every
    shift, near powers of two."""
    out = [2**32 + 1, 2**32 + 15, 2**33 - 9, 10**12 + 39, 2**55 - 55, 2**61 - 1, 2**62 - 57, 2**62 + 135,
           2**63 - 25, 2**63 - 1, 4611686014427387903]
    for _ in range(k):
        b = rng.randrange(33, 64)
        out.append(rng.randrange(2**(b - 1) + 1, 2**b) | 1)
    return [p for p in out if 2**32 < p < 2**63]


def test_medium_fits_and_constants(A):
    """medium_fits holds exactly for 2^32 < p < 2^63 and 1 <= bmax < 2^28;
    M = floor(2^94 / D) for D = p << (s - 1), nd0 = 2^32 - D0, k1 = -(D1 + 1)."""
    rng = random.Random(31)
    out = (C.c_uint32 * 3)()
    for p in _medium_moduli(rng, 1500):
        for b in [0, 1, 2**28 - 1, 2**28, 2**31]:
            assert A.t_medium_fits(p, b) == int(1 <= b < 2**28), (p, b)
        D = p << (64 - p.bit_length() - 1)
        assert 2**62 < D < 2**63
        A.t_medium_consts(p, C.byref(out))
        assert out[0] == 2**94 // D, p
        assert out[1] == (1 << 32) - (D & 0xFFFFFFFF), p
        assert out[2] == (-((D >> 32) + 1)) % (1 << 32), p
    for p in [2**32 - 5, 2**32 + 1 - 2, 2**63 + 29, 2**64 - 59]:
        assert A.t_medium_fits(p, 1000) == int(2**32 < p < 2**63), p


def test_medium_step_against_bigint(A):
    """The medium step maps every x = X << (s - 1) below 2D and n < 2^28 to a
    value below 2D congruent to (X * n mod p) << (s - 1): at the ends of both
    ranges, next to multiples of D, and on chains of random multipliers,
    each against big integers."""
    rng = random.Random(32)
    for p in _medium_moduli(rng, 1500):
        s1 = 64 - p.bit_length() - 1
        D = p << s1
        Xs = [0, 1, 2, p - 1, p, p + 1, 2 * p - 1] + [rng.randrange(2 * p) for _ in range(6)]
        ns = [1, 2, 3, 2**28 - 1, 2**27 + 1, 65536, 1000003] + [rng.randrange(1, 2**28) for _ in range(3)]
        for X in Xs:
            for n in ns:
                got = A.t_medium_walk(p, X << s1, (C.c_uint32 * 1)(n), 1)
                assert got < 2 * D and got % (1 << s1) == 0 and (got >> s1) % p == X * n % p, (p, X, n)
        for _ in range(6):                          # X * n next to a multiple of p
            n = rng.choice([2**28 - 1, rng.randrange(1, 2**28)])
            k = rng.randrange(1, 2 * n)
            for delta in (-1, 0, 1):
                t = k * p + delta
                if t > 0 and t % n == 0 and t // n < 2 * p:
                    X = t // n
                    got = A.t_medium_walk(p, X << s1, (C.c_uint32 * 1)(n), 1)
                    assert got < 2 * D and (got >> s1) % p == t % p, (p, X, n)
        chain = [rng.randrange(1, 2**28) for _ in range(64)]
        X = rng.randrange(2 * p)
        got = A.t_medium_walk(p, X << s1, (C.c_uint32 * 64)(*chain), 64)
        want = X
        for n in chain:
            want = want * n % p
        assert got < 2 * D and (got >> s1) % p == want, (p, X)


def test_strong_prp2_device_arith(A, orc):
    rng = random.Random(5)
    xs = list(range(3, 3000, 2)) + [rng.randrange(3, R) | 1 for _ in range(3000)]
    xs += [(1 << t) + 1 for t in (8, 16, 32, 40, 48, 56)] + [3215031751, 2152302898747, 3474749660383]
    for x in xs:
        assert bool(A.t_strong_prp2(x)) == orc.strong_prp2(x), x


def test_strong_prp2_lockstep_pairs(A, orc):
    """The generator's lockstep test, two values with shared loop bounds (the
    larger bit length and t of the pair): each verdict equals the oracle's
    strong PRP, across bit-length boundaries, large t (N = 2^t * odd + 1),
    Fermat-form N = 2^t + 1, known base-2 strong pseudoprimes and the top of
    the 64-bit range."""
    rng = random.Random(6)
    xs = list(range(3, 2000, 2)) + [rng.randrange(3, R) | 1 for _ in range(2000)]
    xs += [(1 << t) + 1 for t in (8, 16, 32, 40, 48, 56, 63)] + [3215031751, 2152302898747, 3474749660383]
    xs += [(rng.randrange(1, 1 << 20) | 1) << t | 1 for t in (20, 30, 40) for _ in range(20)]
    xs += [R - 59, R - 1, (1 << 63) + 1, (1 << 62) - 57]
    xs = [x for x in xs if 3 <= x < R]
    out = (C.c_int * 2)()
    for k in range(len(xs)):
        a, b = xs[k], xs[rng.randrange(len(xs))]
        A.t_strong_prp2_pair(a, b, out)
        assert (bool(out[0]), bool(out[1])) == (orc.strong_prp2(a), orc.strong_prp2(b)), (a, b)


def test_is_prime_u32(A, orc):
    """The table self-check's 32-bit primality test: a sieve below 2e5, the
    strong pseudoprimes to small base sets (all composite), 2^32 edges, and
    the oracle's 7-base isPrime on random 32-bit integers."""
    import numpy as np
    lim = 200_000
    sv = np.ones(lim, dtype=bool)
    sv[:2] = False
    for i in range(2, int(lim ** 0.5) + 1):
        if sv[i]:
            sv[i * i::i] = False
    for n in range(lim):
        assert bool(A.t_is_prime_u32(n)) == bool(sv[n]), n
    # composites that pass base 2 (2047...), bases 2,3 (1373653), 2,3,5 (25326001),
    # 2,3,5,7 (3215031751), and Carmichael numbers
    for n in [2047, 3277, 4033, 4681, 8321, 1373653, 25326001, 3215031751, 561, 41041, 825265,
              321197185, 4294967295]:
        assert not A.t_is_prime_u32(n), n
    for n in [4294967291, 4294967279, 2147483647, 65537, 61, 7]:
        assert A.t_is_prime_u32(n), n
    rng = random.Random(11)
    for _ in range(20000):
        n = rng.randrange(1 << 32)
        assert bool(A.t_is_prime_u32(n)) == orc.is_prime(n), n


@pytest.mark.parametrize("lo,hi", [(114, 20000), (119, 120), (10**9, 10**9 + 2 * 10**5),
                                   (2**62 - 10**5, 2**62), (R - 10**5, R - 1)])
def test_generate_window_logic_matches_oracle(A, orc, lo, hi):
    """The generator's window arithmetic (sieve masks, range clipping) as the
    kernel lanes run it, against the oracle's candidate set."""
    n = A.t_generate(lo, hi, None, 0)
    out = (C.c_uint64 * max(n, 1))()
    A.t_generate(lo, hi, out, n)
    assert list(out[:n]) == [int(x) for x in orc.candidates(lo, hi)]


@pytest.mark.parametrize("kind,N", [(0, 100), (0, 4999), (0, 99999), (1, 100), (1, 77777), (2, 100), (2, 65432)])
def test_start_residue_against_bigint(L, A, kind, N):
    """The set-up's evaluation of the table in the shifted representation
    gives (X mod p) << s for X = N!, N#, N!/N#, whole and split over 256
    lanes as the table self-check splits it."""
    from pfcsieve_amd.sieve import TableLayout
    t = TableLayout(kind, N)
    ch = (C.c_uint64 * max(len(t.chunks), 1))(*t.chunks)
    gr = (C.c_uint32 * len(t.groups))(*t.groups)
    prim = sympy.primorial(N, nth=False)
    X = [math.factorial(N), prim, math.factorial(N) // prim][kind]
    rng = random.Random(kind * 1000 + N)
    for p in _primes_mix(rng, 40) + [101, 103, 10007, N + 1 if (N + 1) % 2 else N + 2]:
        s = 64 - p.bit_length()
        got = A.t_start_residue(ch, gr, t.npos, t.digit_bits, p)
        assert got == (X % p) << s, (kind, N, p)
    for p in _primes_mix(rng, 5):
        s = 64 - p.bit_length()
        assert A.t_start_residue_split(ch, gr, t.npos, t.digit_bits, p, 256) == (X % p) << s


# ----------------------------------------------------------------- round shares


def _share(A, cost, cap, budget, min_steps):
    n = len(cost)
    arr = C.c_uint64 * max(n, 1)
    out = arr()
    A.t_share_steps(arr(*cost), arr(*cap), n, budget, min_steps, out)
    return list(out)[:n]


def _water_level(demand, budget):
    """Largest integer L with sum(min(d, L)) <= budget (None: every demand fits)."""
    if sum(demand) <= budget:
        return None
    lo, hi = 0, max(demand)
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if sum(min(d, mid) for d in demand) <= budget:
            lo = mid
        else:
            hi = mid - 1
    return lo


def test_round_share_is_water_filling(A):
    """A work-sized round gives every segment the same share of its updates
    and a segment that needs less its whole remainder: the level is the
    largest L with sum(min(cost * cap, L)) <= budget, found here by
    bisection, and a segment below it takes cap steps, the others
    floor(L / cost) (at least min(min_steps, cap))."""
    rng = random.Random(11)
    cases = [([], [], 10, 256),
             ([10**6], [5000], 4 * 10**11, 256),          # alone: budget / cost steps
             ([10**4, 10**8], [10**6, 10**6], 4 * 10**11, 256),
             ([0, 7, 10**9], [300, 10**6, 10**6], 10**9, 256),
             ([1, 1, 1], [10, 10, 10], 10**6, 256)]
    for _ in range(300):
        n = rng.randrange(1, 40)
        cost = [rng.choice([0, rng.randrange(1, 100), rng.randrange(1, 10**9)]) for _ in range(n)]
        cap = [rng.randrange(1, 10**6) for _ in range(n)]
        cases.append((cost, cap, rng.randrange(1, 10**13), rng.choice([1, 256])))
    for cost, cap, budget, m0 in cases:
        got = _share(A, cost, cap, budget, m0)
        demand = [c * k for c, k in zip(cost, cap)]
        L = _water_level(demand, budget)
        want = []
        for c, k, d in zip(cost, cap, demand):
            if L is None or d <= L:
                want.append(k)
            else:
                want.append(min(k, max(L // c, min(m0, k))))
        assert got == want, (cost, cap, budget, m0)
        # the round's updates stay within the budget, apart from the minimum steps
        extra = sum(c * max(0, min(m0, k) - (L // c if c and L is not None else k))
                    for c, k in zip(cost, cap))
        assert sum(c * s for c, s in zip(cost, got)) <= budget + extra


def test_round_share_small_segment_finishes_beside_large_ones(A):
    """A segment of 10^4 candidates beside three of 10^8 gets its whole
    remainder in one round of 4e11 updates; the large ones split the rest
    evenly (the same step count, equal costs)."""
    cost = [10**4, 10**8, 10**8, 10**8]
    cap = [760000] * 4
    got = _share(A, cost, cap, 4 * 10**11, 256)
    assert got[0] == 760000
    assert got[1] == got[2] == got[3] == (4 * 10**11 - 10**4 * 760000) // 3 // 10**8
