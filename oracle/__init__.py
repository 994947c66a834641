"""CPU oracle for the PFCSieve sieve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker.  The product library
(``pfcsieve_amd``) never imports, links or runs anything under ``oracle/``.

The arithmetic lives in ``pfc_oracle.c`` (plain C, unsigned __int128,
OpenMP), built by ``oracle/Makefile`` into ``oracle/_build/liboracle.so``.
Each C function cites the reference file:line it restates.  This module only
binds it with ctypes and returns numpy arrays / Python ints.

``oracle/_ref/`` holds the reference's own self-contained CPU sources
(verifyprime.cpp, putil.c) compiled from /root/reference by ``make ref``;
``ref_is_prime``, ``ref_verify`` and ``ref_parse_uint64`` call them when that
build exists (they are used to pin the restatement, never as the product).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

F, P, Cm = 1, 2, 4          # mode bits (-!, -#, -c)
FACTORIAL, PRIMORIAL, COMPOSITORIAL = 0, 1, 2   # factor record types (cl_sieve.h:4-6)

FACTOR_DTYPE = np.dtype([("p", "<u8"), ("nc", "<i4"), ("type", "<i4")])   # cl_sieve.h:8-12


class OrcFactor(C.Structure):
    _fields_ = [("p", C.c_uint64), ("nc", C.c_int32), ("type", C.c_int32)]


def build(force: bool = False) -> str:
    """Compile the oracle (and oracle/_ref when the reference tree is mounted)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference") and (force or not os.path.isdir(REF_DIR)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64, u32, i32, sz = C.c_uint64, C.c_uint32, C.c_int32, C.c_size_t
        pu64, pu32 = C.POINTER(u64), C.POINTER(u32)
        L.orc_invert.argtypes = [u64]; L.orc_invert.restype = u64
        L.orc_mmul.argtypes = [u64, u64, u64, u64]; L.orc_mmul.restype = u64
        L.orc_madd.argtypes = [u64, u64, u64]; L.orc_madd.restype = u64
        L.orc_strong_prp2.argtypes = [u64]; L.orc_strong_prp2.restype = C.c_int
        L.orc_is_prime.argtypes = [u64]; L.orc_is_prime.restype = C.c_int
        L.orc_verify.argtypes = [u64, u32, i32, i32]; L.orc_verify.restype = C.c_int
        L.orc_wheel_offset.argtypes = [pu64, C.POINTER(i32)]; L.orc_wheel_offset.restype = None
        L.orc_candidates.argtypes = [u64, u64, pu64, sz]; L.orc_candidates.restype = sz
        L.orc_primes.argtypes = [u32, u32, pu32, sz]; L.orc_primes.restype = sz
        L.orc_power_table.argtypes = [u32, pu64, pu32, pu32, sz]; L.orc_power_table.restype = sz
        L.orc_prime_products.argtypes = [u32, pu64, sz]; L.orc_prime_products.restype = sz
        L.orc_comp_products.argtypes = [u32, pu64, sz]; L.orc_comp_products.restype = sz
        L.orc_candidate_record.argtypes = [u32, u64, u32, u32, pu64, pu64]
        L.orc_candidate_record.restype = C.c_int
        L.orc_harvest.argtypes = [C.POINTER(OrcFactor), sz, pu64]; L.orc_harvest.restype = sz
        L.orc_run.argtypes = [u32, u64, u64, u32, u32, C.POINTER(OrcFactor), sz, C.POINTER(sz),
                              C.POINTER(sz), pu64, pu64, pu64, pu64]
        L.orc_run.restype = C.c_int
        L.orc_num_threads.argtypes = []; L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def invert(p): return lib().orc_invert(p)
def mmul(a, b, p, q): return lib().orc_mmul(a, b, p, q)
def madd(a, b, p): return lib().orc_madd(a, b, p)
def strong_prp2(n): return bool(lib().orc_strong_prp2(n))
def is_prime(p): return bool(lib().orc_is_prime(p))
def verify(p, n, c, t): return bool(lib().orc_verify(p, n, c, t))


def wheel_offset(start):
    s = C.c_uint64(start)
    i = C.c_int32(0)
    lib().orc_wheel_offset(C.byref(s), C.byref(i))
    return s.value, i.value


def candidates(pmin, pmax) -> np.ndarray:
    n = lib().orc_candidates(pmin, pmax, None, 0)
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_candidates(pmin, pmax, _p64(out), n)
    return out


def primes(lo, hi) -> np.ndarray:
    n = lib().orc_primes(lo, hi, None, 0)
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_primes(lo, hi, _p32(out), n)
    return out


def power_table(N):
    m = lib().orc_power_table(N, None, None, None, 0)
    prod = np.zeros(m, np.uint64); e = np.zeros(m, np.uint32); b = np.zeros(m, np.uint32)
    lib().orc_power_table(N, _p64(prod), _p32(e), _p32(b), m)
    return prod, e, b


def prime_products(N):
    m = lib().orc_prime_products(N, None, 0)
    out = np.zeros(m, np.uint64)
    lib().orc_prime_products(N, _p64(out), m)
    return out


def comp_products(N):
    m = lib().orc_comp_products(N, None, 0)
    out = np.zeros(m, np.uint64)
    lib().orc_comp_products(N, _p64(out), m)
    return out


def candidate_record(modes, p, nmin, nmax):
    """(ulong8 record after setup+walk, checksum term) for one candidate."""
    rec = np.zeros(8, np.uint64)
    term = C.c_uint64(0)
    rc = lib().orc_candidate_record(modes, p, nmin, nmax, _p64(rec), C.byref(term))
    if rc:
        raise ValueError("invalid job (primorial nmax is prime)")
    return [int(x) for x in rec], term.value


@dataclass
class JobResult:
    checksum: int           # final checksum (device terms + kept factors' n+c)
    dev_checksum: int       # device part alone (check kernels)
    primecount: int
    factorcount: int
    raw_factors: int        # records before the isPrime filter
    factors: np.ndarray     # FACTOR_DTYPE, sorted, filtered


class OracleError(RuntimeError):
    pass


def run(modes, pmin, pmax, nmin, nmax, fcap=10_000_000) -> JobResult:
    fac = (OrcFactor * max(fcap, 1))()
    nraw, nf = C.c_size_t(0), C.c_size_t(0)
    dev, cs, pc, fc = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    rc = lib().orc_run(modes, pmin, pmax, nmin, nmax, fac, fcap, C.byref(nraw), C.byref(nf),
                       C.byref(dev), C.byref(cs), C.byref(pc), C.byref(fc))
    if rc == -1:
        raise OracleError("CPU sieve failure (primorial nmax is prime)")
    if rc:
        raise OracleError(f"oracle run failed rc={rc}")
    arr = np.frombuffer(bytes(fac)[: nf.value * 16], dtype=FACTOR_DTYPE).copy()
    return JobResult(cs.value, dev.value, pc.value, fc.value, nraw.value, arr)


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------------------ reference build


_ref_vp = None
_ref_pu = None


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libverifyprime.so"))


def _ref_lib():
    global _ref_vp, _ref_pu
    if _ref_vp is None:
        _ref_vp = C.CDLL(os.path.join(REF_DIR, "libverifyprime.so"))
        _ref_vp._Z7isPrimem.argtypes = [C.c_uint64]
        _ref_vp._Z7isPrimem.restype = C.c_bool
        _ref_vp._Z6verifymjiiPjm.argtypes = [C.c_uint64, C.c_uint32, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_uint32), C.c_size_t]
        _ref_vp._Z6verifymjiiPjm.restype = C.c_bool
        _ref_vp._Z5m_mulmmmm.argtypes = [C.c_uint64] * 4
        _ref_vp._Z5m_mulmmmm.restype = C.c_uint64
        _ref_vp._Z6invertm.argtypes = [C.c_uint64]
        _ref_vp._Z6invertm.restype = C.c_uint64
        _ref_pu = C.CDLL(os.path.join(REF_DIR, "libputil.so"))
        _ref_pu.parse_uint64.argtypes = [C.POINTER(C.c_uint64), C.c_char_p, C.c_uint64, C.c_uint64]
        _ref_pu.parse_uint64.restype = C.c_int
    return _ref_vp, _ref_pu


def ref_is_prime(p):
    return bool(_ref_lib()[0]._Z7isPrimem(p))


def ref_mmul(a, b, p, q):
    return _ref_lib()[0]._Z5m_mulmmmm(a, b, p, q)


def ref_invert(p):
    return _ref_lib()[0]._Z6invertm(p)


def ref_verify(p, n, c, t, nmax):
    """The reference's verify() with the verifylist its host builds
    (cl_sieve.cpp:1470-1491): primes in [103, nmax] for primorial, composites
    in [45, nmax) for compositorial."""
    if t == PRIMORIAL:
        lst = primes(103, nmax)
    elif t == COMPOSITORIAL:
        pr = set(int(x) for x in primes(45, nmax))
        lst = np.array([n_ for n_ in range(45, nmax) if n_ not in pr], dtype=np.uint32)
    else:
        lst = np.zeros(1, np.uint32)
    lst = np.ascontiguousarray(lst, dtype=np.uint32)
    return bool(_ref_lib()[0]._Z6verifymjiiPjm(p, n, c, t, _p32(lst), len(lst) if t else 0))


def ref_parse_uint64(s: str, lo: int, hi: int):
    out = C.c_uint64(0)
    rc = _ref_lib()[1].parse_uint64(C.byref(out), s.encode(), lo, hi)
    return rc, out.value
