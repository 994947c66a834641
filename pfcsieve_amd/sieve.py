"""Python host API over libpfcsieve.so, mirroring the reference's search driver.

``run_sieve`` is the counterpart of ``cl_sieve(hardware, st, sd)``
(cl_sieve.cpp:1328-1761): it sieves one job and returns the same three
numbers the reference prints and its self-test checks (factorcount,
primecount, checksum; cl_sieve.cpp:1748, 1787-2080), plus the factor lines of
``factors.txt`` (cl_sieve.cpp:565-577, 805-816).  Errors that make the
reference exit raise ``PfcError`` with the reference's message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

from ._native import Factor, JobDesc, JobStats, Tuning, check, lib

MODE_FACTORIAL, MODE_PRIMORIAL, MODE_COMPOSITORIAL = 1, 2, 4
FACTORIAL, PRIMORIAL, COMPOSITORIAL = 0, 1, 2


class Device:
    """One HIP device (replaces the OpenCL platform/context/queue, main.cpp:252-292)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        check(lib().pfc_device_open(ordinal, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self):
        name = C.create_string_buffer(128)
        cus = C.c_int()
        hbm = C.c_uint64()
        check(lib().pfc_device_info(self._h, name, 128, C.byref(cus), C.byref(hbm)))
        return {"name": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}

    def set_tuning(self, primes_per_lane=0, block_size=0, walk_chunk=0, seg_candidates=0, use_graph=0,
                   skip_table_check=0, lds_segment=0, walk_forms_off=0, batch_walk_blocks=0):
        t = Tuning(primes_per_lane, block_size, walk_chunk, seg_candidates, use_graph, skip_table_check,
                   lds_segment, walk_forms_off, batch_walk_blocks)
        check(lib().pfc_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().pfc_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def autotune(self, modes, pmin, pmax, nmin, nmax) -> dict:
        d = job_desc(modes, pmin, pmax, nmin, nmax)
        t = Tuning()
        check(lib().pfc_autotune(self._h, C.byref(d), C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def probe_mul_rates(self):
        r = (C.c_double * 3)()
        check(lib().pfc_probe_mul_rates(self._h, C.byref(r)))
        return {"mad_u64_u32": r[0], "mul_lo_u32": r[1], "mul_hi_u32": r[2]}

    def probe_copy(self, nbytes: int) -> float:
        ms = C.c_double()
        check(lib().pfc_probe_copy(self._h, nbytes, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            lib().pfc_device_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A raw device allocation (pfc_device_malloc) for stage-level calls."""

    def __init__(self, dev: Device, nbytes: int):
        self.dev, self.nbytes = dev, nbytes
        p = C.c_void_p()
        check(lib().pfc_device_malloc(dev.handle, nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, data: bytes):
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().pfc_memcpy(self.dev.handle, self.ptr, buf, len(data), 1))

    def download(self, nbytes: Optional[int] = None) -> bytes:
        n = self.nbytes if nbytes is None else nbytes
        buf = C.create_string_buffer(n)
        check(lib().pfc_memcpy(self.dev.handle, buf, self.ptr, n, 2))
        return buf.raw

    def zero(self):
        self.upload(b"\0" * self.nbytes)

    def free(self):
        if self.ptr:
            lib().pfc_device_free(self.dev.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def job_desc(modes, pmin, pmax, nmin, nmax, factor_capacity=0) -> JobDesc:
    return JobDesc(modes, nmin, nmax, pmin, pmax, factor_capacity)


class Job:
    """One sieve job on a device (workStatus of cl_sieve.h:14-18 plus its buffers)."""

    def __init__(self, dev: Device, modes, pmin, pmax, nmin, nmax, factor_capacity=0):
        self.dev = dev
        self.desc = job_desc(modes, pmin, pmax, nmin, nmax, factor_capacity)
        h = C.c_void_p()
        check(lib().pfc_job_create(dev.handle, C.byref(self.desc), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def run(self, max_segments: int = 0):
        check(lib().pfc_job_run(self._h, max_segments))

    def position(self) -> int:
        p = C.c_uint64()
        check(lib().pfc_job_position(self._h, C.byref(p)))
        return p.value

    def seek(self, p, dev_checksum, primecount):
        check(lib().pfc_job_seek(self._h, p, dev_checksum, primecount))

    def stats(self) -> JobStats:
        s = JobStats()
        check(lib().pfc_job_stats_get(self._h, C.byref(s)))
        return s

    def raw_factors(self):
        n = C.c_size_t()
        check(lib().pfc_job_factors(self._h, None, 0, C.byref(n)))
        arr = (Factor * max(n.value, 1))()
        check(lib().pfc_job_factors(self._h, arr, n.value, C.byref(n)))
        return arr, n.value

    def sync(self):
        check(lib().pfc_job_sync(self._h))

    def result(self):
        """(device checksum part, candidate count, raw factor records sorted) in one call."""
        cs, pc, n = C.c_uint64(), C.c_uint64(), C.c_size_t()
        check(lib().pfc_job_result(self._h, C.byref(cs), C.byref(pc), None, 0, C.byref(n)))
        arr = (Factor * max(n.value, 1))()
        check(lib().pfc_job_result(self._h, C.byref(cs), C.byref(pc), arr, n.value, C.byref(n)))
        return cs.value, pc.value, [(f.p, f.nc, f.type) for f in arr[: n.value]]

    def clear_factors(self):
        check(lib().pfc_job_clear_factors(self._h))

    def destroy(self):
        if self._h:
            lib().pfc_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def footprint(dev: Optional[Device], modes, pmin, pmax, nmin, nmax) -> int:
    d = job_desc(modes, pmin, pmax, nmin, nmax)
    b = C.c_uint64()
    check(lib().pfc_job_footprint(dev.handle if dev else None, C.byref(d), C.byref(b)))
    return b.value


class TableLayout:
    """A start-residue table built on the host (pfc_build_setup_table):
    chunks, group offsets, digit positions and digit width."""

    def __init__(self, kind: int, N: int):
        nch, ng, npos, k = C.c_size_t(), C.c_size_t(), C.c_uint32(), C.c_uint32()
        check(lib().pfc_build_setup_table(kind, N, None, 0, C.byref(nch), None, 0, C.byref(ng), C.byref(npos),
                                          C.byref(k)))
        ch = (C.c_uint64 * max(nch.value, 1))()
        gr = (C.c_uint32 * ng.value)()
        check(lib().pfc_build_setup_table(kind, N, ch, nch.value, C.byref(nch), gr, ng.value, C.byref(ng),
                                          C.byref(npos), C.byref(k)))
        self.kind, self.N = kind, N
        self.chunks = list(ch[: nch.value])
        self.groups = list(gr)
        self.npos, self.digit_bits = npos.value, k.value

    def group(self, b: int, t: int):
        """Chunks of the primes whose exponent digit at position b is t."""
        nd = (1 << self.digit_bits) - 1
        g = (self.npos - 1 - b) * nd + (nd - t)
        return self.chunks[self.groups[g]: self.groups[g + 1]]

    def value(self) -> int:
        """X(N) from the table (big integer), by the set-up's own evaluation
        order: Horner over positions, suffix products over digits."""
        B = 1 << self.digit_bits
        x = 1
        for b in range(self.npos - 1, -1, -1):
            x = x ** B
            S, h = 1, 1
            for t in range(B - 1, 0, -1):
                for c in self.group(b, t):
                    S *= c
                h *= S
            x *= h
        return x

    def products(self) -> int:
        """Modular products one candidate's start residue costs: every chunk,
        the suffix products from each position's largest non-empty digit
        down (none for digit width 1), and the squarings between positions."""
        k, B = self.digit_bits, 1 << self.digit_bits
        n = len(self.chunks) + k * max(self.npos - 1, 0)
        if k > 1:
            for b in range(self.npos):
                top = max((t for t in range(1, B) if self.group(b, t)), default=0)
                n += top
        return n


def setup_table_size(kind: int, N: int):
    """(chunks, digit positions, digit bits) of the start-residue table for
    X(N) with kind 0 = N!, 1 = N#, 2 = N!/N# (host only)."""
    t = TableLayout(kind, N)
    return len(t.chunks), t.npos, t.digit_bits


def setup_table_cost(kind: int, N: int) -> int:
    """Modular products one candidate's start residue X(N) costs in the set-up."""
    return TableLayout(kind, N).products()


def format_factor(f: Factor) -> str:
    buf = C.create_string_buffer(96)
    lib().pfc_format_factor(C.byref(f), buf, 96)
    return buf.value.decode()


def harvest(arr, n, checksum: int, verify_threads: int = 2):
    """getResults' host half (cl_sieve.cpp:496-593): returns (checksum, kept list)."""
    cs = C.c_uint64(checksum)
    kept = C.c_size_t()
    check(lib().pfc_harvest(arr, n, verify_threads, C.byref(cs), C.byref(kept)))
    return cs.value, [arr[i] for i in range(kept.value)]


@dataclass
class SieveResult:
    factorcount: int
    primecount: int
    checksum: int
    lines: List[str] = field(default_factory=list)
    factors: List[tuple] = field(default_factory=list)   # (p, nc, type), sorted
    device_ms: float = 0.0
    segments: int = 0

    def factors_txt(self) -> str:
        """factors.txt content (cl_sieve.cpp:565-577, 805-816)."""
        body = "".join(self.lines)
        if self.factorcount:
            return body + "%016X\n" % self.checksum
        return body + "no factors\n%016X\n" % self.checksum


def run_sieve(modes, pmin, pmax, nmin, nmax, *, device: Optional[Device] = None,
              verify_threads: int = 2, factor_capacity: int = 0) -> SieveResult:
    dev = device or Device(0)
    job = Job(dev, modes, pmin, pmax, nmin, nmax, factor_capacity)
    try:
        job.run()
        st = job.stats()
        arr, n = job.raw_factors()
        checksum, kept = harvest(arr, n, st.dev_checksum, verify_threads)
        res = SieveResult(factorcount=len(kept), primecount=st.primecount, checksum=checksum,
                          device_ms=st.device_ms, segments=st.segments)
        res.lines = [format_factor(f) for f in kept]
        res.factors = [(f.p, f.nc, f.type) for f in kept]
        return res
    finally:
        job.destroy()


class Batcher:
    """Continuous batching of jobs on one device (pfc_batcher_* in include/pfc.h)."""

    def __init__(self, dev: Device):
        self.dev = dev
        h = C.c_void_p()
        check(lib().pfc_batcher_create(dev.handle, C.byref(h)))
        self._h = h
        self._jobs = {}

    def add(self, job: Job):
        check(lib().pfc_batcher_add(self._h, job.handle))
        self._jobs[job.handle.value] = job

    def pending(self) -> int:
        n = C.c_size_t()
        check(lib().pfc_batcher_pending(self._h, C.byref(n)))
        return n.value

    def round(self, max_steps: int = 0) -> List[Job]:
        cap = max(len(self._jobs), 1)
        arr = (C.c_void_p * cap)()
        nd = C.c_size_t()
        check(lib().pfc_batcher_round(self._h, max_steps, arr, cap, C.byref(nd)))
        return [self._jobs.pop(arr[i]) for i in range(min(nd.value, cap))]

    def round_work(self, max_updates: int) -> List[Job]:
        """One round of about max_updates p*n updates, an equal share per
        resident segment (pfc_batcher_round_work)."""
        cap = max(len(self._jobs), 1)
        arr = (C.c_void_p * cap)()
        nd = C.c_size_t()
        check(lib().pfc_batcher_round_work(self._h, max_updates, arr, cap, C.byref(nd)))
        return [self._jobs.pop(arr[i]) for i in range(min(nd.value, cap))]

    def destroy(self):
        if self._h:
            lib().pfc_batcher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def job_result(job: Job, verify_threads: int = 2) -> SieveResult:
    """Harvested result of a finished job (the tail of run_sieve)."""
    st = job.stats()
    arr, n = job.raw_factors()
    checksum, kept = harvest(arr, n, st.dev_checksum, verify_threads)
    res = SieveResult(factorcount=len(kept), primecount=st.primecount, checksum=checksum,
                      device_ms=st.device_ms, segments=st.segments)
    res.lines = [format_factor(f) for f in kept]
    res.factors = [(f.p, f.nc, f.type) for f in kept]
    return res


def run_batch(descs, *, device: Optional[Device] = None, max_steps: int = 0,
              verify_threads: int = 2, progress=None,
              hbm_budget: Optional[int] = None) -> List[SieveResult]:
    """Runs job descriptions (modes, pmin, pmax, nmin, nmax) through one batcher.

    Jobs are admitted in order while the sum of their device footprints
    (pfc_job_footprint) fits hbm_budget (default: 80% of the device's memory);
    finished jobs are harvested and released between rounds and the next ones
    admitted, so a list larger than the device runs in one pass.  Results come
    back in the order of descs.  progress: optional callable receiving one
    line per round."""
    dev = device or Device(0)
    budget = hbm_budget if hbm_budget is not None else int(dev.info()["hbm_bytes"] * 0.8)
    sizes = [footprint(dev, *d) for d in descs]
    for i, size in enumerate(sizes):
        if size > budget:
            raise ValueError(f"job {i} needs {size} bytes of device memory, over the budget of {budget}")
    results: List[Optional[SieveResult]] = [None] * len(descs)
    active = {}             # job handle -> (index, job)
    resident, nxt, rounds = 0, 0, 0
    b = Batcher(dev)
    try:
        while nxt < len(descs) or active:
            while nxt < len(descs) and resident + sizes[nxt] <= budget:
                j = Job(dev, *descs[nxt])
                b.add(j)
                active[j.handle.value] = (nxt, j)
                resident += sizes[nxt]
                nxt += 1
            for job in b.round(max_steps):
                i, j = active.pop(job.handle.value)
                results[i] = job_result(j, verify_threads)
                resident -= sizes[i]
                j.destroy()
            rounds += 1
            if progress:
                progress(f"batch round {rounds}: {len(descs) - nxt + len(active)} of {len(descs)} jobs pending")
        return results
    finally:
        b.destroy()
        for _, j in active.values():
            j.destroy()
